// pib_scalar.h -- scalar transport, one species: the per-mode implicit x
// solve of scalstep (stepthescalar.F90:67-241) on the device, FP64.
//
// Each (ky, kz) Fourier mode owns one lane.  The lane walks x three times:
// a forward Thomas sweep that forms the right-hand side of phistep on the
// fly from three rolling phif planes (:115-140), the back substitution, and
// the Sherman-Morrison correction of the x wrap with the heat-ratio twist
// phif(0) = heat_ratio * phif(nx), phif(nx+1) = phif(1) / heat_ratio_new
// (:142-196).  Fields are plane-major [i][kz][ky] with nx+1 planes, so the
// lanes of a wavefront touch consecutive modes of one plane (coalesced).
// A workgroup is one wavefront: a 128^3 grid has only 8320 modes, and
// one-wave workgroups spread them over 130 CUs instead of 33.  Each sweep
// loads PHI_K planes into registers before it runs their PHI_K steps of the
// recurrence, so a lane keeps PHI_K plane loads in flight.  Each row takes
// one reciprocal of its pivot, and the multiplier of the next row and the
// three right-hand sides multiply by it; the correction factor (v.y)/(1+v.z)
// is formed once.  The reference divides in each of these places; the
// divisions are a lane's longest dependent chains.  This rounds differently
// from the reference by about an ulp.
#pragma once

#define PHI_T 64
#define PHI_K 16

struct ScalArgs {
  int nx, my, mz, my2, rks;
  double gamma;
  const double* ratio;        // device: heat_ratio, heat_ratio_new (read at the kernel's start)
  const DevState* st;
  C2<double>* phi;            // [i][kz][ky], i = 0..nx
  const C2<double>* nlphi;    // nlphif, onlphif, ffphi: same layout, i < nx read
  const C2<double>* onlphi;
  const C2<double>* ffphi;
  double* scr;                // four vectors [s][i][mode]: gam, y_re, y_aux, y_im
};

__global__ void __launch_bounds__(PHI_T) k_phistep(ScalArgs A, int* status) {
  const int E = A.my2 * A.mz;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= E) return;
  const int nx = A.nx, ky = q % A.my2, kz = q / A.my2;
  const DevState* st = A.st;
  const int r = A.rks - 1;
  const double c1 = st->coef[r][0], c2 = st->coef[r][1], c3 = st->coef[r][2], c4 = st->coef[r][3];
  const double dx = st->dx, dx2 = dx * dx, dt = st->s.dt, g = A.gamma;
  // w2 = wy^2 + wz^2 of this mode (initialize_flo.F90 wavenumbers)
  const double twopi = 2.0 * 3.14159265358979323846264338327950288;
  const double wy = twopi * (double)ky / ((double)A.my * dx);
  const double wz = kz < A.mz / 2 ? twopi * (double)kz / ((double)A.mz * dx)
                                  : -(twopi * (double)(A.mz - kz) / ((double)A.mz * dx));
  const double w2 = wy * wy + wz * wz;
  // tridiagonal rows (:118-124, 133-134) and Sherman-Morrison vectors (:178-183)
  const double a = -(dt * g * c2 / dx2), c = a;
  const double b = 1.0 + dt * g * c2 * (2.0 / dx2 + w2);
  const double HR = A.ratio[0], HRN = A.ratio[1];
  const double alpha = HRN, beta = 1.0 / HRN;
  const double a1 = a * HRN, cn = c / HRN;
  const double v1 = cn / beta, vn = a1 / alpha;
  const double b0 = b - alpha * v1, bn = b - beta * vn;
  const double f0 = 1.0 - dt * c1 * g * w2, dc3 = dt * c3, dc4 = dt * c4, dg1 = dt * g * c1;
  const double gl = dg1 / dx2;  // weight of the explicit second difference
  const size_t ms = (size_t)E;
  const C2<double>* ph = A.phi + q;
  const C2<double>* __restrict__ nl = A.nlphi + q;
  const C2<double>* __restrict__ on = A.onlphi + q;
  const C2<double>* __restrict__ fp = A.ffphi + q;
  double* __restrict__ gs = A.scr + q;
  double* __restrict__ y1s = A.scr + (size_t)nx * ms + q;
  double* __restrict__ y2s = A.scr + (size_t)2 * nx * ms + q;
  double* __restrict__ y3s = A.scr + (size_t)3 * nx * ms + q;
  auto rhs = [&](int i, C2<double> pm, C2<double> p0, C2<double> pp, C2<double> n, C2<double> w,
                 C2<double> f) {
    double rx = f0 * p0.x, ry = f0 * p0.y;
    rx = (rx + dc3 * n.x) + dt * f.x;
    ry = (ry + dc3 * n.y) + dt * f.y;
    rx = rx + dc4 * w.x;
    ry = ry + dc4 * w.y;
    double lx, ly;
    if (i == 0) {                       // x-periodic first row (:136-140)
      lx = (pp.x - 2.0 * p0.x) + HR * pm.x;
      ly = (pp.y - 2.0 * p0.y) + HR * pm.y;
    } else {
      lx = (pm.x + pp.x) - 2.0 * p0.x;
      ly = (pm.y + pp.y) - 2.0 * p0.y;
    }
    return mk(rx + gl * lx, ry + gl * ly);
  };
  int status_ = 0;
  // forward sweep (tridag3 of the rhs pair and auxu)
  C2<double> pm = ph[(size_t)(nx - 1) * ms], p0 = ph[0], pp = ph[ms];
  double bet = b0;
  if (bet == 0.0) status_ = PIB_ERR_PIVOT;
  C2<double> rv = rhs(0, pm, p0, pp, nl[0], on[0], fp[0]);
  double ib = 1.0 / bet;  // the one division a row: every pivot use multiplies by it
  double y1 = rv.x * ib, y2 = alpha * ib, y3 = rv.y * ib;
  y1s[0] = y1;
  y2s[0] = y2;
  y3s[0] = y3;
  gs[0] = 0.0;
  for (int i0 = 1; i0 < nx; i0 += PHI_K) {
    C2<double> P[PHI_K], N[PHI_K], W[PHI_K], F[PHI_K];
#pragma unroll
    for (int u = 0; u < PHI_K; ++u) {
      const int i = i0 + u;
      if (i < nx) {
        const size_t o = (size_t)i * ms;
        P[u] = ph[o + ms];
        N[u] = nl[o];
        W[u] = on[o];
        F[u] = fp[o];
      }
    }
#pragma unroll
    for (int u = 0; u < PHI_K; ++u) {
      const int i = i0 + u;
      if (i < nx) {
        pm = p0;
        p0 = pp;
        pp = P[u];
        const double gam = c * ib;
        bet = (i == nx - 1 ? bn : b) - a * gam;
        if (bet == 0.0) { bet = 1.0e-30; status_ = PIB_ERR_PIVOT; }
        rv = rhs(i, pm, p0, pp, N[u], W[u], F[u]);
        const double u2 = i == nx - 1 ? beta : 0.0;
        ib = 1.0 / bet;
        y1 = (rv.x - a * y1) * ib;
        y2 = (u2 - a * y2) * ib;
        y3 = (rv.y - a * y3) * ib;
        const size_t o = (size_t)i * ms;
        gs[o] = gam;
        y1s[o] = y1;
        y2s[o] = y2;
        y3s[o] = y3;
      }
    }
  }
  // back substitution; y1..y3 hold row nx-1 already
  const double z_n = y2, yr_n = y1, yi_n = y3;
  for (int i0 = nx - 2; i0 >= 0; i0 -= PHI_K) {
    double G[PHI_K], Y1[PHI_K], Y2[PHI_K], Y3[PHI_K];
#pragma unroll
    for (int u = 0; u < PHI_K; ++u) {
      const int i = i0 - u;
      if (i >= 0) {
        const size_t o = (size_t)i * ms;
        G[u] = gs[o + ms];
        Y1[u] = y1s[o];
        Y2[u] = y2s[o];
        Y3[u] = y3s[o];
      }
    }
#pragma unroll
    for (int u = 0; u < PHI_K; ++u) {
      const int i = i0 - u;
      if (i >= 0) {
        const size_t o = (size_t)i * ms;
        y1 = Y1[u] - G[u] * y1;
        y2 = Y2[u] - G[u] * y2;
        y3 = Y3[u] - G[u] * y3;
        y1s[o] = y1;
        y2s[o] = y2;
        y3s[o] = y3;
      }
    }
  }
  // Sherman-Morrison correction and the twisted wrap plane (:185-196)
  const double vdotz = v1 * y2 + vn * z_n;
  const double vdotyr = v1 * y1 + vn * yr_n;
  const double vdotyi = v1 * y3 + vn * yi_n;
  const double den = 1.0 + vdotz;
  const double sr = vdotyr / den, si = vdotyi / den;
  C2<double>* pw = A.phi + q;
  C2<double> first = mk(y1 - sr * y2, y3 - si * y2);
  pw[0] = first;
  for (int i0 = 1; i0 < nx; i0 += PHI_K) {
    double Z[PHI_K], Y1[PHI_K], Y3[PHI_K];
#pragma unroll
    for (int u = 0; u < PHI_K; ++u) {
      const size_t o = (size_t)(i0 + u) * ms;
      if (i0 + u < nx) {
        Z[u] = y2s[o];
        Y1[u] = y1s[o];
        Y3[u] = y3s[o];
      }
    }
#pragma unroll
    for (int u = 0; u < PHI_K; ++u)
      if (i0 + u < nx)
        pw[(size_t)(i0 + u) * ms] = mk(Y1[u] - sr * Z[u], Y3[u] - si * Z[u]);
  }
  pw[(size_t)nx * ms] = mk(first.x / HRN, first.y / HRN);
  // raised into the context's device status word, kept if already raised
  if (status_) atomicCAS(status, 0, status_);
}

// scalstep's heat-ratio roll (stepthescalar.F90:89-90) on the device ratios
// (heat_ratio, heat_ratio_new, heat_ratio_old)
__global__ void k_ratio_roll(double* r) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    r[2] = r[0];
    r[0] = r[1];
  }
}
