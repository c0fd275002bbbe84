// tw_io.cpp -- reference-format front-end: problem files, source/receiver
// lists, model files, shot setup and the text result files.  Host-only code
// on top of the public C-ABI (include/tw.h); nothing here touches the device
// directly.
//
// Formats (reference tree):
//   InputFile.txt  cuda/src/cpp/input_processing.cpp:11-137 (keys model, grids,
//                  dxM, dxS, bdry, Nt, MAXdt, CFL, frequency, delay, device;
//                  values '_'-separated after the key, anything after them on
//                  the line is ignored)
//   options        input_processing.cpp:140-200
//   SrcRcv.txt     cuda/src/cpp/inverse.cpp:133-348
//   model files    inverse.cpp:6-130 (raw FP64, x outer)
//   results        cuda/src/cpp/forward.cpp:215-327, folder
//                  cuda/src/cuda_main/nvcc_device_FDM_2D.cu:755-772
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <fstream>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "tw_internal.h"

namespace {

int err(int code, const std::string& msg) { return tw::host_fail(code, msg.c_str()); }

// '_'-separated tokens of a line: the key and the values that follow it.
// A value ends at the next '_' or at white space (the rest of the line, a
// comment in the reference's files, is not part of it).
std::vector<std::string> tokens(const std::string& line) {
  std::vector<std::string> t;
  size_t b = 0;
  while (true) {
    size_t e = line.find('_', b);
    std::string tok = line.substr(b, e == std::string::npos ? std::string::npos : e - b);
    size_t ws = tok.find_first_of(" \t\r\n/");
    if (ws != std::string::npos) {
      t.push_back(tok.substr(0, ws));
      break;
    }
    t.push_back(tok);
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return t;
}

bool to_i64(const std::string& v, int64_t* out) {
  if (v.empty()) return false;
  char* end = nullptr;
  errno = 0;
  long long x = strtoll(v.c_str(), &end, 10);
  if (errno || *end) return false;
  *out = x;
  return true;
}

bool to_i32(const std::string& v, int32_t* out) {
  int64_t x;
  if (!to_i64(v, &x) || x < INT32_MIN || x > INT32_MAX) return false;
  *out = (int32_t)x;
  return true;
}

bool to_f64(const std::string& v, double* out) {
  if (v.empty()) return false;
  char* end = nullptr;
  errno = 0;
  double x = strtod(v.c_str(), &end);
  if (errno || *end) return false;
  *out = x;
  return true;
}

// Boundary letters as written in the files: XL XR YL YR (physical X first).
// tw_config.bc is xL xR yL yR in data axes, physical X = data y.
bool set_bdry(tw_input* in, const std::string& v) {
  if (v.size() < 4) return false;
  char u[4];
  for (int i = 0; i < 4; i++) {
    u[i] = (char)toupper((unsigned char)v[i]);
    if (u[i] != 'F' && u[i] != 'P' && u[i] != 'C') return false;
  }
  in->bc[0] = u[2];
  in->bc[1] = u[3];
  in->bc[2] = u[0];
  in->bc[3] = u[1];
  return true;
}

// Grid type letters as written in the files, physical X first -> data x, y.
bool set_grid(const std::string& v, char g[2]) {
  if (v.size() != 2) return false;
  for (int i = 0; i < 2; i++)
    if (v[i] != 'N' && v[i] != 'M') return false;
  g[1] = v[0];
  g[0] = v[1];
  return true;
}

int grid_id(const char g[2]) {
  if (g[0] == 'N') return g[1] == 'N' ? TW_GRID_NN : TW_GRID_NM;
  return g[1] == 'N' ? TW_GRID_MN : TW_GRID_MM;
}

// Fields of one grid in the reference's N_soln order.
std::vector<int32_t> grid_fields(int physics, int grid) {
  switch (grid) {
    case TW_GRID_NN: return physics == TW_ELASTIC ? std::vector<int32_t>{TW_SXY} : std::vector<int32_t>{};
    case TW_GRID_MN: return {TW_VY};
    case TW_GRID_NM: return {TW_VX};
    default: return physics == TW_ELASTIC ? std::vector<int32_t>{TW_SXX, TW_SYY} : std::vector<int32_t>{TW_P};
  }
}

struct Shot {
  const tw_srcrcv_entry* src = nullptr;
  int src_ix = 0, src_iy = 0;
  std::vector<const tw_srcrcv_entry*> rcv;
  std::vector<int32_t> rcv_ix, rcv_iy;
  std::vector<int32_t> field, fx, fy;  // registration order of tw_set_receivers
};

int locate(const tw_config* cfg, const tw_srcrcv_entry* e, int32_t* ix, int32_t* iy) {
  return tw_locate(cfg, grid_id(e->grid), e->num[0], e->den[0], e->num[1], e->den[1], ix, iy);
}

int build_shot(const tw_config* cfg, const tw_srcrcv_entry* entries, int32_t n, int32_t src_index, Shot* sh) {
  if (!entries || n < 0) return err(TW_EINVAL, "bad source/receiver list");
  for (int i = 0; i < n; i++)
    if (entries[i].kind == 0 && entries[i].src_index == src_index) sh->src = &entries[i];
  if (!sh->src) return err(TW_EINVAL, "source index " + std::to_string(src_index) + " not in the list");
  int e = locate(cfg, sh->src, &sh->src_ix, &sh->src_iy);
  if (e) return e;
  for (int i = 0; i < n; i++) {
    const tw_srcrcv_entry* r = &entries[i];
    if (r->kind != 1 || r->src_index != src_index) continue;
    int32_t ix, iy;
    if ((e = locate(cfg, r, &ix, &iy))) return e;
    const std::vector<int32_t> fs = grid_fields(cfg->physics, grid_id(r->grid));
    if (fs.empty()) return err(TW_EINVAL, "receiver on a grid this physics does not have");
    sh->rcv.push_back(r);
    sh->rcv_ix.push_back(ix);
    sh->rcv_iy.push_back(iy);
    for (int32_t f : fs) {
      sh->field.push_back(f);
      sh->fx.push_back(ix);
      sh->fy.push_back(iy);
    }
  }
  if (sh->rcv.empty()) return err(TW_EINVAL, "no receiver found for source " + std::to_string(src_index));
  return TW_OK;
}

// Physical X first: the reference prints grid_type.at(1), grid_type.at(0)
// and the locations iy, ix (forward.cpp:239-251).
std::string tag(const char g[2], int ix, int iy) {
  std::string s;
  s += g[1];
  s += g[0];
  return s + "_" + std::to_string(iy) + "_" + std::to_string(ix);
}

}  // namespace

extern "C" {

void tw_input_default(tw_input* in) {
  if (!in) return;
  memset(in, 0, sizeof *in);
  in->model_m[0] = in->model_m[1] = -1;
  in->grid_m[0] = in->grid_m[1] = -1;
  in->num_dx_model = in->num_dx = -1;
  in->den_dx_model = in->den_dx = 1;
  memcpy(in->bc, "FFFF", 4);
  in->nt = -1;
  in->dt_max = (double)(1 << 30);
  in->cfl = 0.625;
  in->central_f = 5.0;
  in->time_delay = 0.25;
  in->device = -1;
  in->energy = 'Y';
  snprintf(in->medium, sizeof in->medium, "%s", "homogeneous");
}

int tw_read_input_file(const char* path, tw_input* in) {
  if (!path || !in) return err(TW_EINVAL, "null argument");
  std::ifstream f(path);
  if (!f.is_open()) return err(TW_EINVAL, std::string("unable to open file: ") + path);
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    lineno++;
    const std::vector<std::string> t = tokens(line);
    const std::string& key = t[0];
    auto need = [&](size_t k) { return t.size() >= k + 1; };
    auto bad = [&]() { return err(TW_EINVAL, std::string(path) + ":" + std::to_string(lineno) + ": bad value for " + key); };
    // model_<X>_<Y>, grids_<X>_<Y>: physical X is data axis 1
    if (key == "model" || key == "grids") {
      int32_t X, Y;
      if (!need(2) || !to_i32(t[1], &X) || !to_i32(t[2], &Y)) return bad();
      int32_t* m = key == "model" ? in->model_m : in->grid_m;
      m[1] = X;
      m[0] = Y;
    } else if (key == "dxM" || key == "dxS") {
      int64_t a, b;
      if (!need(2) || !to_i64(t[1], &a) || !to_i64(t[2], &b)) return bad();
      if (key == "dxM") {
        in->num_dx_model = a;
        in->den_dx_model = b;
      } else {
        in->num_dx = a;
        in->den_dx = b;
      }
    } else if (key == "bdry") {
      if (!need(1) || !set_bdry(in, t[1])) return bad();
    } else if (key == "Nt") {
      if (!need(1) || !to_i32(t[1], &in->nt)) return bad();
    } else if (key == "MAXdt") {
      if (!need(1) || !to_f64(t[1], &in->dt_max)) return bad();
    } else if (key == "CFL") {
      if (!need(1) || !to_f64(t[1], &in->cfl)) return bad();
    } else if (key == "frequency") {
      if (!need(1) || !to_f64(t[1], &in->central_f)) return bad();
    } else if (key == "delay") {
      if (!need(1) || !to_f64(t[1], &in->time_delay)) return bad();
    } else if (key == "device") {
      if (!need(1) || !to_i32(t[1], &in->device)) return bad();
    }
  }
  return TW_OK;
}

int tw_input_option(tw_input* in, const char* name, const char* value) {
  if (!in || !name || !value) return err(TW_EINVAL, "null argument");
  const std::string n(name), v(value);
  bool ok = true;
  if (n == "medium") {
    if (v.size() >= sizeof in->medium) return err(TW_EINVAL, "medium name too long");
    snprintf(in->medium, sizeof in->medium, "%s", value);
  } else if (n == "B_type") {
    ok = set_bdry(in, v);
  } else if (n == "Nt") {
    ok = to_i32(v, &in->nt);
  } else if (n == "dt_max") {
    ok = to_f64(v, &in->dt_max);
  } else if (n == "CFL") {
    ok = to_f64(v, &in->cfl);
  } else if (n == "energy") {
    ok = !v.empty();
    if (ok) in->energy = (char)toupper((unsigned char)v[0]);
  } else if (n == "f") {
    ok = to_f64(v, &in->central_f);
  } else if (n == "t0") {
    ok = to_f64(v, &in->time_delay);
  } else if (n == "device") {
    ok = to_i32(v, &in->device);
  } else {
    return err(TW_EINVAL, "unknown option " + n);
  }
  return ok ? TW_OK : err(TW_EINVAL, "bad value for option " + n + ": " + v);
}

int tw_input_to_config(const tw_input* in, int32_t physics, int32_t precision, int32_t cpst, tw_config* cfg) {
  if (!in || !cfg) return err(TW_EINVAL, "null argument");
  for (int d = 0; d < 2; d++) {
    if (in->grid_m[d] <= 0) return err(TW_EINVAL, "simulation size not set (grids_)");
    if (in->num_dx <= 0 || in->den_dx <= 0) return err(TW_EINVAL, "simulation spacing not set (dxS_)");
    // input_processing.cpp:206-210: model and simulation extents agree
    if (in->model_m[d] > 0 &&
        (int64_t)in->model_m[d] * in->num_dx_model * in->den_dx != (int64_t)in->grid_m[d] * in->num_dx * in->den_dx_model)
      return err(TW_EINVAL, "model size needs to match simulation size on direction " + std::to_string(d));
  }
  if (in->energy != 'Y' && in->energy != 'N') return err(TW_EINVAL, "energy can only be N or Y");
  if (in->nt <= 0) return err(TW_EINVAL, "Nt not set");
  tw_config c;
  tw_config_default(&c);
  c.physics = physics;
  c.mx = in->grid_m[0];
  c.my = in->grid_m[1];
  c.num_dx = in->num_dx;
  c.den_dx = in->den_dx;
  c.dx = (double)in->num_dx / (double)in->den_dx;
  c.dt = in->dt_max * in->cfl;
  memcpy(c.bc, in->bc, 4);
  c.precision = precision;
  c.cpst = cpst;
  c.nt_max = in->nt;
  c.record_energy = in->energy == 'Y';
  c.device = in->device < 0 ? 0 : in->device;
  if (tw_config_check(&c)) return TW_EINVAL;  // message set by the check
  *cfg = c;
  return TW_OK;
}

int tw_read_srcrcv_file(const char* path, tw_srcrcv_entry* out, int32_t cap, int32_t* n) {
  if (!path || !n) return err(TW_EINVAL, "null argument");
  std::ifstream f(path);
  if (!f.is_open()) return err(TW_EINVAL, std::string("unable to open file: ") + path);
  std::vector<tw_srcrcv_entry> v;
  std::set<int32_t> src_seen;
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    lineno++;
    const std::vector<std::string> t = tokens(line);
    const bool src = strcasecmp(t[0].c_str(), "Src") == 0, rcv = strcasecmp(t[0].c_str(), "Rcv") == 0;
    if (!src && !rcv) continue;
    auto bad = [&]() { return err(TW_EINVAL, std::string(path) + ":" + std::to_string(lineno) + ": malformed line"); };
    tw_srcrcv_entry e;
    memset(&e, 0, sizeof e);
    e.kind = rcv;
    e.rcv_index = -1;
    size_t k = 1;
    if (t.size() < (rcv ? 8u : 7u)) return bad();
    if (!to_i32(t[k++], &e.src_index)) return bad();
    if (rcv) {
      if (!src_seen.count(e.src_index))
        return err(TW_EINVAL, "receiver (source index " + std::to_string(e.src_index) + ") before its source");
      if (!to_i32(t[k++], &e.rcv_index)) return bad();
    }
    if (!set_grid(t[k++], e.grid)) return bad();
    // locations: physical X (num, den) then Y (num, den) -> data y, then x
    int64_t q[4];
    for (int i = 0; i < 4; i++)
      if (!to_i64(t[k++], &q[i])) return bad();
    e.num[1] = q[0];
    e.den[1] = q[1];
    e.num[0] = q[2];
    e.den[0] = q[3];
    if (src) src_seen.insert(e.src_index);
    v.push_back(e);
  }
  // duplicate checks of input_SrcRcv_locations (inverse.cpp:277-347)
  auto key = [](const tw_srcrcv_entry& e) {
    return std::make_pair(std::string(e.grid, 2),
                          std::to_string(e.num[0]) + "/" + std::to_string(e.den[0]) + "," + std::to_string(e.num[1]) + "/" +
                              std::to_string(e.den[1]));
  };
  std::set<std::pair<std::string, std::string>> sl;
  std::set<int32_t> si;
  for (const auto& e : v) {
    if (e.kind) continue;
    if (!sl.insert(key(e)).second) return err(TW_EINVAL, "duplicate source location (source " + std::to_string(e.src_index) + ")");
    if (!si.insert(e.src_index).second) return err(TW_EINVAL, "duplicate source index " + std::to_string(e.src_index));
  }
  for (int32_t s : si) {
    std::set<std::pair<std::string, std::string>> rl;
    std::set<int32_t> ri;
    for (const auto& e : v) {
      if (!e.kind || e.src_index != s) continue;
      if (!rl.insert(key(e)).second)
        return err(TW_EINVAL, "duplicate receiver location (source " + std::to_string(s) + ", receiver " + std::to_string(e.rcv_index) + ")");
      if (!ri.insert(e.rcv_index).second)
        return err(TW_EINVAL, "duplicate receiver index " + std::to_string(e.rcv_index) + " for source " + std::to_string(s));
    }
  }
  *n = (int32_t)v.size();
  if (!out) return TW_OK;
  if ((int32_t)v.size() > cap) return err(TW_ERANGE, "more entries than the output holds");
  for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
  return TW_OK;
}

int tw_read_model_file(const char* path, int32_t mx_model, int32_t my_model, const char bc[4], double* out) {
  if (!path || !out || mx_model <= 0 || my_model <= 0) return err(TW_EINVAL, "bad argument");
  const int64_t nx = (int64_t)mx_model + 1, ny = (int64_t)my_model + 1;
  FILE* f = fopen(path, "rb");
  if (!f) return err(TW_EINVAL, std::string("unable to open file: ") + path);
  fseek(f, 0, SEEK_END);
  const long bytes = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (bytes != (long)(nx * ny * 8)) {
    fclose(f);
    return err(TW_EINVAL, std::string("size of ") + path + " not as expected: " + std::to_string(bytes) + " bytes for " +
                              std::to_string(nx * ny) + " model points");
  }
  const size_t got = fread(out, 8, (size_t)(nx * ny), f);
  fclose(f);
  if (got != (size_t)(nx * ny)) return err(TW_EINVAL, std::string("short read: ") + path);
  // periodic compatibility (inverse.cpp:100-127)
  if (bc && bc[0] == 'P' && bc[1] == 'P')
    for (int64_t iy = 0; iy < ny; iy++)
      if (out[iy] != out[(nx - 1) * ny + iy]) return err(TW_EINVAL, "model not compatible with periodic x");
  if (bc && bc[2] == 'P' && bc[3] == 'P')
    for (int64_t ix = 0; ix < nx; ix++)
      if (out[ix * ny] != out[ix * ny + ny - 1]) return err(TW_EINVAL, "model not compatible with periodic y");
  return TW_OK;
}

int tw_setup_shot(tw_solver* s, const tw_input* in, const tw_srcrcv_entry* entries, int32_t n, int32_t src_index) {
  if (!s || !in) return err(TW_EINVAL, "null argument");
  tw_config cfg;
  int e = tw_get_config(s, &cfg);
  if (e) return e;
  Shot sh;
  if ((e = build_shot(&cfg, entries, n, src_index, &sh))) return e;
  if ((e = tw_set_source_ricker(s, grid_id(sh.src->grid), sh.src_ix, sh.src_iy, in->central_f, in->time_delay))) return e;
  return tw_set_receivers(s, (int32_t)sh.field.size(), sh.field.data(), sh.fx.data(), sh.fy.data());
}

int tw_write_results(tw_solver* s, const char* folder, const tw_srcrcv_entry* entries, int32_t n, int32_t src_index) {
  if (!s || !folder) return err(TW_EINVAL, "null argument");
  tw_config cfg;
  int e = tw_get_config(s, &cfg);
  if (e) return e;
  Shot sh;
  if ((e = build_shot(&cfg, entries, n, src_index, &sh))) return e;
  // the solver must record exactly this shot's receivers, in the order the
  // files are named after (tw_setup_shot registers them so)
  int32_t nreg = 0;
  if ((e = tw_get_receivers(s, 0, &nreg, nullptr, nullptr, nullptr))) return e;
  if (nreg != (int32_t)sh.field.size()) return err(TW_ESTATE, "the solver records other receivers than this shot's");
  std::vector<int32_t> rf(nreg), rx(nreg), ry(nreg);
  if (nreg > 0 && (e = tw_get_receivers(s, nreg, &nreg, rf.data(), rx.data(), ry.data()))) return e;
  if (rf != sh.field || rx != sh.fx || ry != sh.fy) return err(TW_ESTATE, "the solver records other receivers than this shot's");
  int32_t it = 0;
  if ((e = tw_current_step(s, &it))) return e;
  const int nt = it < cfg.nt_max ? it : cfg.nt_max;
  std::vector<double> rec(sh.field.size() * (size_t)(nt > 0 ? nt : 1));
  int32_t got = 0;
  if ((e = tw_read_receivers(s, rec.data(), &got))) return e;
  if (got != nt) return err(TW_ESTATE, "receiver records do not match the step count");
  const double sc = cfg.source_scaling;
  const std::string stag = std::string(folder) + "/Signal_src_" + std::to_string(src_index) + "_" + tag(sh.src->grid, sh.src_ix, sh.src_iy);
  size_t row = 0;
  for (size_t r = 0; r < sh.rcv.size(); r++) {
    const int nf = (int)grid_fields(cfg.physics, grid_id(sh.rcv[r]->grid)).size();
    const std::string path = stag + "_rcv_" + std::to_string(sh.rcv[r]->rcv_index) + "_" + tag(sh.rcv[r]->grid, sh.rcv_ix[r], sh.rcv_iy[r]) + ".txt";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return err(TW_EINVAL, "cannot write " + path);
    for (int t = 0; t < nt; t++) {
      fprintf(f, "% 6d  ", t);
      for (int k = 0; k < nf; k++) fprintf(f, "% 16.15e  ", sc * rec[(row + k) * (size_t)nt + t]);
      fprintf(f, "\n");
    }
    fclose(f);
    row += nf;
  }
  if (cfg.record_energy) {
    std::vector<double> ek(nt > 0 ? nt : 1), e0(ek.size()), e1(ek.size());
    if ((e = tw_read_energy(s, ek.data(), e0.data(), e1.data(), &got))) return e;
    const std::string path = std::string(folder) + "/Energy_src_" + std::to_string(src_index) + "_" + tag(sh.src->grid, sh.src_ix, sh.src_iy) + ".txt";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return err(TW_EINVAL, "cannot write " + path);
    for (int t = 0; t < got; t++) fprintf(f, "%16.15e\n", sc * sc * (ek[t] + (e0[t] + e1[t]) / 2.));
    fclose(f);
  }
  return TW_OK;
}

int tw_write_snapshots(tw_solver* s, const char* folder) {
  if (!s || !folder) return err(TW_EINVAL, "null argument");
  tw_config cfg;
  int e = tw_get_config(s, &cfg);
  if (e) return e;
  // grids in the reference's order NN, MN, NM, MM; names "Sxy", "Vx", "Vy",
  // "SMM" (data axes, grid.hpp:155-159); the file name gives the type with
  // physical X first; elastic fields per grid in soln order, acoustic as the
  // single field of its grid
  static const char* kName[4] = {"Sxy", "Vx", "Vy", "SMM"};
  static const char kType[4][2] = {{'N', 'N'}, {'M', 'N'}, {'N', 'M'}, {'M', 'M'}};
  for (int g = 0; g < 4; g++) {
    const std::vector<int32_t> fs = grid_fields(cfg.physics, g);
    for (size_t i = 0; i < fs.size(); i++) {
      int32_t rows = 0, cols = 0;
      if ((e = tw_field_shape(&cfg, fs[i], &rows, &cols))) return e;
      std::vector<double> v((size_t)rows * cols);
      if ((e = tw_get_field(s, fs[i], v.data()))) return e;
      const std::string path = std::string(folder) + "/Snapshot_" + kType[g][1] + kType[g][0] + "_" + kName[g] + "_" +
                               std::to_string(i) + ".txt";
      FILE* f = fopen(path.c_str(), "w");
      if (!f) return err(TW_EINVAL, "cannot write " + path);
      for (size_t k = 0; k < v.size(); k++) fprintf(f, "% 16.15e\n", cfg.source_scaling * v[k]);
      fclose(f);
    }
  }
  return TW_OK;
}

int tw_result_folder(const tw_input* in, const tw_config* cfg, const char* root, char* out, int32_t cap) {
  if (!in || !cfg || !root || !out || cap <= 0) return err(TW_EINVAL, "bad argument");
  static const char* prec[4] = {"fp64", "fp32", "bf16", "fp16"};
  static const char* cp[10] = {"0_0", "3_R", "-3_R", "6_R", "3_C", "-3_C", "6_C", "K", "K8", "K8S"};
  if (cfg->precision < 0 || cfg->precision > 3 || cfg->cpst < 0 || cfg->cpst > TW_CPST_K8S)
    return err(TW_EINVAL, "bad configuration");
  char dt[64];
  snprintf(dt, sizeof dt, "dt_%16.15e", cfg->dt);
  dt[4] = 'p';  // the reference replaces the decimal point (nvcc_device_FDM_2D.cu:755-756)
  const std::string p = std::string(root) + "/" + (cfg->physics == TW_ELASTIC ? "elastic" : "acoustic") + "/" + in->medium +
                        "/device/grid_size_" + std::to_string(in->num_dx) + "_" + std::to_string(in->den_dx) + "/Nt_" +
                        std::to_string(in->nt) + "/" + prec[cfg->precision] + "/" + dt + "/" + cp[cfg->cpst];
  if ((int32_t)p.size() + 1 > cap) return err(TW_ERANGE, "folder name longer than the output");
  memcpy(out, p.c_str(), p.size() + 1);
  return TW_OK;
}

}  // extern "C"
