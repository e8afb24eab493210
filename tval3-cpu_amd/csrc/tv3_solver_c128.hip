// tv3_solver_c128.hip — the complex<double> solver: the TVAL3 inner loop with
// complex double vectors and double scalars (the reference's
// mat<complex<double>> overloads, 2d/tval3_cpu_2d.cpp:37-41, 64-68, 91-95;
// 3d/tval3_cpu_3d.cpp:65, 103, 141).  The solver source is compiled again
// with its vector type set to cdouble; this unit exports tv3_solve_c128,
// tv3_solve_batch_c128 and tv3_gemv_c128 and runs W = 1 groups only.
#define TV3_VEC tv3::cdouble
#define TV3_SOLVER_VARIANT 3
#include "tv3_solver.hip"
