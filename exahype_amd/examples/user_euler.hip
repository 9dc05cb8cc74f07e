// user_euler.hip -- a user's PDE compiled into the device kernels
// (exahype_amd/user_solver.hpp), as a user of the library would write it: the
// compressible Euler equations with the reference interface's eigenvalues
// vector, the reflecting-wall boundaryValues of the reference's Euler
// demonstrator (Demonstrators/EulerADERDG/MyEulerSolver_ADERDG.cpp:63-88), and
// an adjustPointSolution that sets an entropy wave at t = 0.  Built into
// exahype_amd/libexa_user_example.so; the tests drive it through
// exa_grid_create_user.
#include <exahype_amd/user_solver.hpp>

namespace {

constexpr double kGamma = 1.4;

// Entropy wave rho = 1 + 0.2 sin(2 pi (x + y + z)), u = v = w = 1, p = 1.
__device__ inline void entropy_wave(const double* x, double* Q) {
  const double twopi = 6.283185307179586476925286766559;
  const double rho = 1.0 + 0.2 * sin(twopi * (1.0 * x[0] + 1.0 * x[1] + 1.0 * x[2]));
  Q[0] = rho;
  Q[1] = rho * 1.0;
  Q[2] = rho * 1.0;
  Q[3] = rho * 1.0;
  Q[4] = 1.0 / (kGamma - 1.0) + 0.5 * rho * (1.0 * 1.0 + 1.0 * 1.0 + 1.0 * 1.0);
}

struct UserEuler {
  static constexpr int V = 5;
  static constexpr int P = 0;
  static constexpr double CFL = 0.9;

  template <typename T>
  __device__ static T pressure(const T* Q, T irho) {
    return (T(kGamma) - T(1)) * (Q[4] - T(0.5) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]) * irho);
  }

  template <typename T>
  __device__ static void flux(const T* Q, T (*F)[V]) {
    const T irho = T(1.0) / Q[0];
    const T p = pressure(Q, irho);
    for (int d = 0; d < 3; ++d) {
      F[d][0] = Q[d + 1];
      for (int m = 1; m < 4; ++m) F[d][m] = m == d + 1 ? irho * Q[d + 1] * Q[m] + p : irho * Q[d + 1] * Q[m];
      F[d][4] = irho * Q[d + 1] * (Q[4] + p);
    }
  }

  template <typename T>
  __device__ static void eigenvalues(const T* Q, int d, T* lambda) {
    const T irho = T(1.0) / Q[0];
    const T p = pressure(Q, irho);
    const T u_n = Q[d + 1] * irho;
    const T c = sqrt(T(kGamma) * p * irho);
    lambda[0] = u_n - c;
    lambda[1] = u_n;
    lambda[2] = u_n;
    lambda[3] = u_n;
    lambda[4] = u_n + c;
  }

  template <typename T>
  __device__ static void algebraic_source(const T*, T* S) {
    for (int m = 0; m < V; ++m) S[m] = T(0);
  }

  template <typename T>
  __device__ static void ncp(const T*, const T (*)[V], T* BgradQ) {
    for (int m = 0; m < V; ++m) BgradQ[m] = T(0);
  }

  // reflecting wall: the inside state with the normal momentum negated, and
  // its normal flux
  template <typename T>
  __device__ static void boundary_values(const double*, double, double, int, int normal, const T*, const T* Qin,
                                         T* Fout, T* Qout) {
    for (int m = 0; m < V; ++m) Qout[m] = Qin[m];
    Qout[1 + normal] = -Qout[1 + normal];
    T F[3][V];
    flux(Qout, F);
    for (int m = 0; m < V; ++m) Fout[m] = F[normal][m];
  }

  // initial condition at t = 0; afterwards a relaxation towards it,
  // Q += dt / 2 (Q_0(x) - Q) (a deterministic use of x, t and dt)
  __device__ static void adjust_point_solution(const double* x, double t, double dt, double* Q) {
    double Q0[V];
    entropy_wave(x, Q0);
    if (t == 0.0) {
      for (int m = 0; m < V; ++m) Q[m] = Q0[m];
    } else {
      for (int m = 0; m < V; ++m) Q[m] += 0.5 * dt * (Q0[m] - Q[m]);
    }
  }
};

}  // namespace

EXA_USER_SOLVER(exa_user_euler_order3, UserEuler, 3, exa::kFlux)
EXA_USER_SOLVER(exa_user_euler_order5, UserEuler, 5, exa::kFlux)
