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
#include <vector>

#include "pib_kernels.h"

// Node-sorted (point, corner) records of one stencil of the reversal pairs:
// target node t of ntarget sums the records toff[t]..toff[t+1]-1, record r
// being corner weight rw[r] of point rpt[r] (a point = body*nbnd + l).
struct AdjRecords {
  int ntarget;
  const int *tnode, *toff, *rpt;
  const double* rw;
};
enum { ADJ_REC_INNER = 0, ADJ_REC_OUTER = 1, ADJ_REC_PRES = 2 };

template <typename T>
struct AdjArgs {
  C2<T> *au, *aunew, *ap, *apnew, *anl, *anlp, *aff, *Z;  // spectral, device layout
  // physical: aubcp, aom, apbcp, appr adjoints of ubcp, omega, pbcp, ppr;
  // afr of fr; S the spread of the reversal forcing's adjoint at the inner
  // points (the adjoints of nlbc, onlbc, diffn are -c3 S, -c4 S, -c12 S and
  // the inner-point share of the adjoint of ubcp is cf S); Q the eight
  // transposed transforms of form_nl
  T *aubcp, *aom, *afr, *apbcp, *appr, *S, *Q;
  double* sc;            // [0..2] adjoint of umean, [3..5] dJ/dmpg, [6..8] frmean_bar/N
  double* axc;           // dJ/dxc (nbody, 3) column-major
  double* vs;            // transposed line solves: 12 lines of nx doubles per mode
  double* part;          // 3*nx per-plane sums
  const double* wforce;  // this step's force weights (nbody, 3), or null
  // the reversal forcing's adjoint per point, (h[3], cf kappa P h[3]), and
  // the records that spread it at the inner, outer and pressure points
  double* coef;
  AdjRecords rec[3];
};

struct AdjGeom {
  int sh;          // compiled plane shape (0: generic)
  int nt_plane;    // lanes of a generic plane workgroup
  size_t lds;      // dynamic LDS of a plane workgroup
  int nbody, nbnd;
  hipStream_t stream;
  double c12;      // c1 + c2 of the step's stage
  bool capturing;  // enqueued under graph capture: kernel attributes already set
};

// The kernels of one adjoint step, enqueued on g.stream.  The forward
// fields of the step (ubcp, pbcp, nlbc, onlbc, diffn, omega, ppr, mask) must
// be current in A; on return aunew/apnew/anlp hold the adjoint of the step's
// inputs u, p, nl (the caller swaps them with au/ap/anl).
// Builds the three record sets of the current bed's active reversal pairs
// (device memory from hipMalloc, listed in mem for the caller to free);
// waits on the stream.
void adjoint_build_records(const Particles& P, int nx, int my, int mz, hipStream_t s, AdjRecords rec[3],
                           std::vector<void*>& mem);

template <typename T>
void adjoint_step_kernels(const KArgs<T>& A, const Particles& P, const AdjArgs<T>& J, const AdjGeom& g);
