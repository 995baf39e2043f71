// drivers.h -- the four outer-loop algorithms with the reference's class
// names and constructors (SIMPLEC.h:92, SIMPLER.h:99, SIMPLERR.h:100,
// IDEAL.h:100).  Solve() calls NavierStokesBase methods in the reference's
// order; the fields are device-resident in NSBase.ctx.
#pragma once

#include <array>
#include <ctime>
#include <vector>

#include "NSBase.h"

namespace ns3d_host {

// One RES row: it Q v_moy RMR K gradP normdiv reldiv velocityNorm maxdiv
// it_u it_v it_w it_pc.
using ResRow = std::array<double, 14>;

struct RunLog {
  std::vector<ResRow> rows;
  int iterations = 0;
  bool converged = false;
  double K = 0, Q = 0, v_moy = 0, RMR = 0, normdiv = 0, reldiv = 0, maxdiv = 0;
  double seconds = 0;
};

class SIMPLECSOLVER {
 public:
  int save_step = 20, maxit;
  double simpletol;
  time_t start, end;
  RunLog log;
  SIMPLECSOLVER(double simpletol, int maxit);
  void Solve(NavierStokesBase& NSBase);
};

class SIMPLERSOLVER {
 public:
  int save_step = 20, N1, maxit;
  double simpletol;
  time_t start, end;
  RunLog log;
  SIMPLERSOLVER(double simpletol, int maxit, int N1 = 1);
  void Solve(NavierStokesBase& NSBase);
};

class SIMPLERRSOLVER {
 public:
  int save_step = 20, N1, maxit;
  double simpletol;
  time_t start, end;
  RunLog log;
  SIMPLERRSOLVER(double simpletol, int maxit, int N1);
  void Solve(NavierStokesBase& NSBase);
};

class IDEALSOLVER {
 public:
  int save_step = 20, N1, N2, maxit;
  double simpletol;
  time_t start, end;
  RunLog log;
  IDEALSOLVER(double simpletol, int maxit, int N1 = 2, int N2 = 2);
  void Solve(NavierStokesBase& NSBase);
};

}  // namespace ns3d_host
