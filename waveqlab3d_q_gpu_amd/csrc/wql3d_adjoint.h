// Discrete-adjoint launchers (wql3d_adjoint.hip) used by the C-ABI layer.
#pragma once

#include "wql3d_internal.h"

namespace wql {

// Adjoint state of one block: Fb / Eb adjoint of (F, eta), DFb / DEb adjoint
// of the rate registers (DF, Deta), g the gradient planes (rho, lambda, mu),
// Z 18 planes of workspace.  Same plane stride as the block.
template <typename T>
struct AdjArgs {
  double* Fb;
  T* Eb;
  double* DFb;
  T* DEb;
  double* g[3];
  T* Z;
  int xcd;   // workgroup order of the per-node kernels over the XCDs (xcd_logical)
};

// reverse of one fused stage whose input state is the block's current F / eta
template <typename T, typename St>
hipError_t launch_adjoint_stage(const StepArgs<T, St>& a, const AdjArgs<T>& j, int mesh, double A,
                                double Bdt, hipStream_t s);
// Fb at receiver nodes idx[r] += g[r][0..8]
hipError_t launch_adjoint_inject(double* Fb, long long P, const long long* idx, int n,
                                 const double* g, hipStream_t s);

}  // namespace wql
