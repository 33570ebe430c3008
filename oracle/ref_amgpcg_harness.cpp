// ref_amgpcg_harness.cpp -- C-ABI shim over the reference's standalone
// amgpcg_solve (test/solver/test_amgpcg.cpp:29-70), which is a static function
// of that test program: the file is compiled in place through #include, its
// main() renamed.  Separate translation unit because test/test_common.h
// declares its own `struct Grid`.  Test infrastructure only (oracle/_ref).
#define main reference_test_amgpcg_main
#include REF_AMGPCG_TEST
#undef main

extern "C" int ref_amgpcg(int N, const double* b, double* x, int max_iter, double tol, int* iters) {
    Grid gb(N);
    for (size_t c = 0; c < gb.v.size(); c++) gb.v[c] = b[c];
    int it = 0;
    Grid gx = amgpcg_solve(gb, max_iter, tol, it);
    for (size_t c = 0; c < gx.v.size(); c++) x[c] = gx.v[c];
    *iters = it;
    return 0;
}
