// tv3_solver_f64.hip — the double/double solver: the TVAL3 inner loop with
// FP64 vectors (T_comp = T_scal = double, the reference's mat<double>
// overloads, 2d/tval3_cpu_2d.cpp:26-29, 3d/tval3_cpu_3d.cpp:47, 88, 124).
// The solver source is compiled a second time with its vector type set to
// double; this unit exports tv3_solve_f64, tv3_solve_batch_f64 and
// tv3_gemv_f64 and runs W = 1 groups only.
#define TV3_VEC double
#define TV3_SOLVER_VARIANT 1
#include "tv3_solver.hip"
