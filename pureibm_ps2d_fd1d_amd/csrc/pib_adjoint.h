// pib_adjoint.h -- interface between the context (pib_gpu.hip) and the
// kernels of the adjoint step (pib_adjoint.hip).
//
// The reference has no adjoint.  The adjoint step is the discrete transpose
// of one fixed-bed AB2/CN step of this library with respect to u, p, nl,
// umean, the particle centres and mpg (see include/pureibm_gpu.h).  The
// transposes used:
//   c2r (row-column convention)  x = Re sum c_ky Z e^{+i theta}, c_ky = 1 at
//     ky = 0 and my/2, else 2        ->  Z_bar = c_ky*my*mz * r2c(x_bar)
//   r2c  Y = (my*mz)^-1 sum x e^{-i theta}  ->  x_bar = c2r(Y_bar/c_ky)/(my*mz)
//   multiplication by (i k) -> by (-i k); the symmetric Toeplitz line matrices
//   are their own transposes; trilinear interpolation <-> spreading with the
//   same weights, plus the derivative of the weights in the position.
#pragma once
#include "pib_kernels.h"

template <typename T>
struct AdjArgs {
  C2<T> *au, *aunew, *ap, *apnew, *anl, *anlp, *aff, *Z;  // spectral, device layout
  T *aubcp, *anlbc, *aonlbc, *adiffn, *aom, *afr, *apbcp, *appr, *Q;  // physical
  double* sc;            // [0..2] adjoint of umean, [3..5] dJ/dmpg, [6..8] frmean_bar/N
  double* axc;           // dJ/dxc (nbody, 3) column-major
  double* vs;            // transposed line solves: 12 lines of nx doubles per mode
  double* part;          // 3*nx per-plane sums
  const double* wforce;  // this step's force weights (nbody, 3), or null
};

struct AdjGeom {
  int sh;          // compiled plane shape (0: generic)
  int nt_plane;    // lanes of a generic plane workgroup
  size_t lds;      // dynamic LDS of a plane workgroup
  int nbody, nbnd;
  hipStream_t stream;
};

// The kernels of one adjoint step, enqueued on g.stream.  The forward
// fields of the step (ubcp, pbcp, nlbc, onlbc, diffn, omega, ppr, mask) must
// be current in A; on return aunew/apnew/anlp hold the adjoint of the step's
// inputs u, p, nl (the caller swaps them with au/ap/anl).
template <typename T>
void adjoint_step_kernels(const KArgs<T>& A, const Particles& P, const AdjArgs<T>& J, const AdjGeom& g);
