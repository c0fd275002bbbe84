// Command-line driver for the reference's single-core acoustic solver.
//
// TEST INFRASTRUCTURE ONLY (builds the oracle's golden vectors; the product
// never uses it).  The reference's single-core program
// (singlecore/src/main_cpp/cpu_acoustic.cpp) contains a command-line parser,
// namespace_inputs::command_line_input_processing
// (singlecore/src/cpp/namespaces.hpp:679-730: -Nt, -multiple, -cpst, ...),
// but its main() never calls it, so it always runs its defaults (Nt = 60000,
// no compensation).  This file compiles that program unchanged, under another
// name for its main(), and calls the reference's own parser first, so the
// options take effect exactly as in the reference's MPI build
// (mpi/src/main_cpp/cpu_acoustic.cpp:41 calls the same parser).  Built by
// oracle/ref_build.sh from the sources where they lie.
#define main tw_reference_singlecore_main
#include "cpu_acoustic.cpp"
#undef main

int main(int argc, char* argv[]) {
  namespace_inputs::command_line_input_processing(argc, argv);
  return tw_reference_singlecore_main(argc, argv);
}
