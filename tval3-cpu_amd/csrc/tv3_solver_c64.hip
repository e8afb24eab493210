// tv3_solver_c64.hip — the complex<float> solver: the TVAL3 inner loop with
// complex float vectors and float scalars (the reference's
// mat<complex<float>> overloads, 2d/tval3_cpu_2d.cpp:31-35, 58-62, 85-89;
// 3d/tval3_cpu_3d.cpp:56, 97, 132).  The solver source is compiled again
// with its vector type set to cfloat; this unit exports tv3_solve_c64,
// tv3_solve_batch_c64 and tv3_gemv_c64 and runs W = 1 groups only.
#define TV3_VEC tv3::cfloat
#define TV3_SOLVER_VARIANT 2
#include "tv3_solver.hip"
