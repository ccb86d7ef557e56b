/* C restatement of the reference Ros2 column step - TEST INFRASTRUCTURE ONLY.
 *
 * Used as the timed CPU baseline of bench.py and as a second check of the
 * NumPy oracle (oracle/ros2_np.py).  Follows the reference's own dataflow:
 *   chemistry rhs / Jacobian  chem.py:83-214 (sums in ascending row/slot order)
 *   diffusion                 jax_step.py:67-246
 *   block Thomas              solver.py:33-112: per block a partial-pivot LU
 *                             (getrf, first-max pivot), inverse of the
 *                             previous block by lu_solve(., I), Schur update
 *   Ros2 combine              jax_step.py:249-315
 * The atmosphere comes packed in the record layout of vk_atm_record_layout
 * (include/vulcan_ros2.h).  Nothing in the product links this file.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NAKB (6.02214086e23 * 1.38064852e-16)

typedef struct {
  int nz, ni, S, T;
  const int64_t *r_idx, *p_idx;
  const double *r_st, *p_st;
  const uint8_t* is3;
} Net;

typedef struct {
  const double *Kzz, *dzi, *vz, *Hpi, *Ti, *Tco, *g, *M;
  const double *ms, *alpha, *top_flux, *bot_flux, *bot_vdep, *gas;
  const double *Dzz, *vm, *vs, *flags;
} Atm;

static Atm atm_view(const double* b, int nz, int ni) {
  Atm a;
  int64_t p = 0, nzm = nz - 1;
  a.Kzz = b + p; p += nzm; a.dzi = b + p; p += nzm; a.vz = b + p; p += nzm;
  a.Hpi = b + p; p += nzm; a.Ti = b + p; p += nzm; a.Tco = b + p; p += nz;
  a.g = b + p; p += nz; a.M = b + p; p += nz; a.ms = b + p; p += ni;
  a.alpha = b + p; p += ni; a.top_flux = b + p; p += ni; a.bot_flux = b + p; p += ni;
  a.bot_vdep = b + p; p += ni; a.gas = b + p; p += ni; a.Dzz = b + p; p += nzm * ni;
  a.vm = b + p; p += (int64_t)nz * ni; a.vs = b + p; p += nzm * ni; a.flags = b + p;
  return a;
}

static double ipow(double y, double nu) {
  int n = (int)nu;
  if (n <= 0) return 1.0;
  double r = y;
  for (int i = 1; i < n; ++i) r *= y;
  return r;
}

static double posp(double v) { return v > 0 ? v : 0; }
static double negp(double v) { return v < 0 ? v : 0; }

/* rates and their slot derivatives at one layer; rate[S], dr[S*T] */
static void layer_rates(const Net* n, const double* yz, double Mz, const double* k, int z,
                        double* rate, double* dr) {
  const int T = n->T;
  for (int r = 0; r < n->S; ++r) {
    double fac[4], pr = 1.0;
    const double kr = k[(int64_t)r * n->nz + z];
    for (int s = 0; s < T; ++s) {
      const int64_t sp = n->r_idx[r * T + s];
      const double st = n->r_st[r * T + s];
      const double yv = sp < n->ni ? yz[sp] : 1.0;
      fac[s] = st > 0 ? ipow(yv, st) : 1.0;
      pr = pr * fac[s];
    }
    double rt = kr * pr;
    if (n->is3[r]) rt = rt * Mz;
    rate[r] = rt;
    if (dr) {
      for (int s = 0; s < T; ++s) {
        const int64_t sp = n->r_idx[r * T + s];
        const double st = n->r_st[r * T + s];
        const double yv = sp < n->ni ? yz[sp] : 1.0;
        double lv = 1.0;
        for (int l = 0; l < T; ++l)
          if (l != s) lv = lv * fac[l];
        const double pm1 = st > 0 ? ipow(yv, st - 1.0) : 0.0;
        double d = st * pm1 * lv * kr;
        if (n->is3[r]) d = d * Mz;
        dr[r * T + s] = d;
      }
    }
  }
}

static void chem_layer(const Net* n, const double* rate, double* out, double* loss, double* prod) {
  const int ni = n->ni, T = n->T;
  memset(loss, 0, sizeof(double) * (ni + 1));
  memset(prod, 0, sizeof(double) * (ni + 1));
  for (int r = 0; r < n->S; ++r)
    for (int s = 0; s < T; ++s) {
      loss[n->r_idx[r * T + s]] += n->r_st[r * T + s] * rate[r];
      prod[n->p_idx[r * T + s]] += n->p_st[r * T + s] * rate[r];
    }
  for (int i = 0; i < ni; ++i) out[i] = prod[i] - loss[i];
}

static void jac_layer(const Net* n, const double* dr, double* J /* (ni+1)^2 scratch */) {
  const int ni = n->ni, T = n->T, W = ni + 1;
  memset(J, 0, sizeof(double) * W * W);
  for (int r = 0; r < n->S; ++r)
    for (int so = 0; so < 2 * T; ++so) {
      const int64_t oi = so < T ? n->r_idx[r * T + so] : n->p_idx[r * T + so - T];
      const double oc = so < T ? -n->r_st[r * T + so] : n->p_st[r * T + so - T];
      for (int sj = 0; sj < T; ++sj) {
        const int64_t ii = n->r_idx[r * T + sj];
        J[oi * W + ii] += oc * dr[r * T + sj];
      }
    }
}

static double grav_fac(const Atm* v, int i, int a, int b) {
  return -1.0 / v->Hpi[a] + v->ms[i] * v->g[b] / (NAKB * v->Ti[a]) +
         v->alpha[i] / v->Ti[a] * (v->Tco[a + 1] - v->Tco[a]) / v->dzi[a];
}

/* total diffusion coefficients At, Bt, Ct of species i at layer z */
static void diff_coef(const Atm* v, int nz, int ni, int z, int i, double ysm, double ys0,
                      double ysp, double* At, double* Bt, double* Ct) {
  const double* dzi = v->dzi;
  const double* Kzz = v->Kzz;
  const double* vz = v->vz;
  const double use_vm = v->flags[0] != 0, use_set = v->flags[1] != 0;
  double A, B, C, Ai, Bi, Ci;
  if (z == 0) {
    A = -1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysp + ys0) / 2.0 / ys0 - posp(vz[0]) / dzi[0];
    B = 1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysp + ys0) / 2.0 / ysp - negp(vz[0]) / dzi[0];
    C = 0.0;
    const double D0 = v->Dzz[i];
    const double g0 = 1.0 / dzi[0] * D0 / 2.0 * grav_fac(v, i, 0, 0);
    const double bA = (1.0 - use_vm) * g0 + use_vm * (-posp(v->vm[i]) / dzi[0]) + use_set * (-posp(v->vs[i]) / dzi[0]);
    const double bB = (1.0 - use_vm) * g0 + use_vm * (-negp(v->vm[i]) / dzi[0]) + use_set * (-negp(v->vs[i]) / dzi[0]);
    Ai = -1.0 / dzi[0] * (D0 / dzi[0]) * (ysp + ys0) / 2.0 / ys0 + bA;
    Bi = 1.0 / dzi[0] * (D0 / dzi[0]) * (ysp + ys0) / 2.0 / ysp + bB;
    Ci = 0.0;
  } else if (z == nz - 1) {
    const int t = nz - 2;
    A = -1.0 / dzi[t] * (Kzz[t] / dzi[t]) * (ys0 + ysm) / 2.0 / ys0 + negp(vz[t]) / dzi[t];
    C = 1.0 / dzi[t] * (Kzz[t] / dzi[t]) * (ys0 + ysm) / 2.0 / ysm + posp(vz[t]) / dzi[t];
    B = 0.0;
    const double Dt = v->Dzz[(int64_t)t * ni + i];
    const double gt = -1.0 / dzi[t] * Dt / 2.0 *
                      (-1.0 / v->Hpi[t] + v->ms[i] * v->g[nz - 1] / (NAKB * v->Ti[t]) +
                       v->alpha[i] / v->Ti[t] * (v->Tco[nz - 1] - v->Tco[nz - 2]) / dzi[t]);
    const double vmt = v->vm[(int64_t)(nz - 1) * ni + i], vst = v->vs[(int64_t)t * ni + i];
    const double bA = (1.0 - use_vm) * gt + use_vm * (negp(vmt) / dzi[t]) + use_set * (negp(vst) / dzi[t]);
    const double bC = (1.0 - use_vm) * gt + use_vm * (posp(vmt) / dzi[t]) + use_set * (posp(vst) / dzi[t]);
    Ai = -1.0 / dzi[t] * (Dt / dzi[t]) * (ys0 + ysm) / 2.0 / ys0 + bA;
    Ci = 1.0 / dzi[t] * (Dt / dzi[t]) * (ys0 + ysm) / 2.0 / ysm + bC;
    Bi = 0.0;
  } else {
    const double dza = 0.5 * (dzi[z - 1] + dzi[z]);
    A = -1.0 / dza * (Kzz[z] / dzi[z] * (ysp + ys0) / 2.0 + Kzz[z - 1] / dzi[z - 1] * (ys0 + ysm) / 2.0) / ys0 -
        (posp(vz[z]) - negp(vz[z - 1])) / dza;
    B = 1.0 / dza * Kzz[z] / dzi[z] * (ysp + ys0) / 2.0 / ysp - negp(vz[z]) / dza;
    C = 1.0 / dza * Kzz[z - 1] / dzi[z - 1] * (ys0 + ysm) / 2.0 / ysm + posp(vz[z - 1]) / dza;
    const double Dj = v->Dzz[(int64_t)z * ni + i], Djm = v->Dzz[(int64_t)(z - 1) * ni + i];
    const double inv2 = 1.0 / (2.0 * dza);
    const double gA = inv2 * (Dj * grav_fac(v, i, z, z) - Djm * grav_fac(v, i, z - 1, z));
    const double gB = inv2 * Dj * grav_fac(v, i, z, z + 1);
    const double gC = -inv2 * Djm * grav_fac(v, i, z - 1, z - 1);
    const double vmj = v->vm[(int64_t)z * ni + i], vmm = v->vm[(int64_t)(z - 1) * ni + i];
    const double vsj = v->vs[(int64_t)z * ni + i], vsm = v->vs[(int64_t)(z - 1) * ni + i];
    const double Ax = ((1.0 - use_vm) * gA + use_vm * (-(posp(vmj) - negp(vmm)) / dza)) + use_set * (-(posp(vsj) - negp(vsm)) / dza);
    const double Bx = ((1.0 - use_vm) * gB + use_vm * (-negp(vmj) / dza)) + use_set * (-negp(vsj) / dza);
    const double Cx = ((1.0 - use_vm) * gC + use_vm * (posp(vmm) / dza)) + use_set * (posp(vsm) / dza);
    Ai = -1.0 / dza * (Dj / dzi[z] * (ysp + ys0) / 2.0 + Djm / dzi[z - 1] * (ys0 + ysm) / 2.0) / ys0 + Ax;
    Bi = 1.0 / dza * Dj / dzi[z] * (ysp + ys0) / 2.0 / ysp + Bx;
    Ci = 1.0 / dza * Djm / dzi[z - 1] * (ys0 + ysm) / 2.0 / ysm + Cx;
  }
  *At = A + Ai;
  *Bt = B + Bi;
  *Ct = C + Ci;
}

static double ysum_gas(const Atm* v, const double* yz, int ni) {
  double s = 0;
  for (int i = 0; i < ni; ++i)
    if (v->gas[i] != 0) s += yz[i];
  return s;
}

/* f = chem + diffusion at y; optionally the LHS pieces (diag blocks without
 * the c0/Jacobian part: dd, sup, sub). */
static void total_rhs(const Net* n, const Atm* v, const double* y, const double* k, double* f,
                      double* dd, double* sup, double* sub, double* rate, double* loss,
                      double* prod) {
  const int nz = n->nz, ni = n->ni;
  for (int z = 0; z < nz; ++z) {
    layer_rates(n, y + (int64_t)z * ni, v->M[z], k, z, rate, NULL);
    chem_layer(n, rate, f + (int64_t)z * ni, loss, prod);
  }
  for (int z = 0; z < nz; ++z) {
    const double ysm = z > 0 ? ysum_gas(v, y + (int64_t)(z - 1) * ni, ni) : 1.0;
    const double ys0 = ysum_gas(v, y + (int64_t)z * ni, ni);
    const double ysp = z < nz - 1 ? ysum_gas(v, y + (int64_t)(z + 1) * ni, ni) : 1.0;
    for (int i = 0; i < ni; ++i) {
      double At, Bt, Ct;
      diff_coef(v, nz, ni, z, i, ysm, ys0, ysp, &At, &Bt, &Ct);
      const double yi = y[(int64_t)z * ni + i];
      double d;
      if (z == 0) d = At * yi + Bt * y[(int64_t)ni + i];
      else if (z == nz - 1) d = At * yi + Ct * y[(int64_t)(z - 1) * ni + i];
      else d = At * yi + Bt * y[(int64_t)(z + 1) * ni + i] + Ct * y[(int64_t)(z - 1) * ni + i];
      if (z == nz - 1 && v->flags[2] != 0) d = d + v->top_flux[i] / v->dzi[nz - 2];
      if (z == 0 && v->flags[3] != 0) d = d + (v->bot_flux[i] - yi * v->bot_vdep[i]) / v->dzi[0];
      f[(int64_t)z * ni + i] += d;
      if (dd) {
        double dv = At;
        if (z == 0 && v->flags[3] != 0) dv = dv + (-v->bot_vdep[i] / v->dzi[0]);
        dd[(int64_t)z * ni + i] = dv;
        if (z < nz - 1) sup[(int64_t)z * ni + i] = Bt;
        if (z > 0) sub[(int64_t)(z - 1) * ni + i] = Ct;
      }
    }
  }
}

/* getrf with partial pivoting (first max), column-major-free row-major form */
static void lu_factor(double* a, int n, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(a[(int64_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(a[(int64_t)i * n + k]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        const double t = a[(int64_t)k * n + j];
        a[(int64_t)k * n + j] = a[(int64_t)p * n + j];
        a[(int64_t)p * n + j] = t;
      }
    const double rp = 1.0 / a[(int64_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = a[(int64_t)i * n + k] * rp;
      a[(int64_t)i * n + k] = l;
      if (l != 0.0) {
        double* ri = a + (int64_t)i * n;
        const double* rk = a + (int64_t)k * n;
        for (int j = k + 1; j < n; ++j) ri[j] -= l * rk[j];
      }
    }
  }
}

static void lu_solve(const double* a, int n, const int* piv, double* b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) { const double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
  for (int i = 1; i < n; ++i) {
    double s = b[i];
    const double* ri = a + (int64_t)i * n;
    for (int j = 0; j < i; ++j) s -= ri[j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    const double* ri = a + (int64_t)i * n;
    for (int j = i + 1; j < n; ++j) s -= ri[j] * b[j];
    b[i] = s / ri[i];
  }
}

/* inverse of an LU-factored block, rows of inv = solves against unit vectors */
static void lu_inverse(const double* a, int n, const int* piv, double* inv, double* col) {
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (i == c);
    lu_solve(a, n, piv, col);
    for (int i = 0; i < n; ++i) inv[(int64_t)i * n + c] = col[i];
  }
}

typedef struct {
  double *lu, *inv, *col, *rmod, *f, *dd, *sup, *sub, *k1, *yk2, *rhs2, *k2, *rate, *dr, *J,
      *loss, *prod;
  int* piv;
} Work;

static int work_alloc(Work* w, int nz, int ni, int S, int T) {
  const int64_t blk = (int64_t)ni * ni, st = (int64_t)nz * ni;
  w->lu = malloc(sizeof(double) * nz * blk);
  w->inv = malloc(sizeof(double) * blk);
  w->col = malloc(sizeof(double) * ni);
  w->piv = malloc(sizeof(int) * nz * ni);
  double** v[] = {&w->rmod, &w->f, &w->dd, &w->sup, &w->sub, &w->k1, &w->yk2, &w->rhs2, &w->k2};
  for (unsigned q = 0; q < sizeof(v) / sizeof(v[0]); ++q) *v[q] = malloc(sizeof(double) * st);
  w->rate = malloc(sizeof(double) * S);
  w->dr = malloc(sizeof(double) * S * T);
  w->J = malloc(sizeof(double) * (ni + 1) * (ni + 1));
  w->loss = malloc(sizeof(double) * (ni + 1));
  w->prod = malloc(sizeof(double) * (ni + 1));
  return w->lu && w->inv && w->piv && w->J ? 0 : -1;
}

static void work_free(Work* w) {
  void* v[] = {w->lu, w->inv, w->col, w->piv, w->rmod, w->f, w->dd, w->sup, w->sub, w->k1,
               w->yk2, w->rhs2, w->k2, w->rate, w->dr, w->J, w->loss, w->prod};
  for (unsigned q = 0; q < sizeof(v) / sizeof(v[0]); ++q) free(v[q]);
}

static void bt_solve(const Work* w, int nz, int ni, const double* sup_neg, const double* sub_neg,
                     const double* rhs, double* x, double* rmod, double* tmp) {
  memcpy(rmod, rhs, sizeof(double) * ni);
  for (int j = 1; j < nz; ++j) {
    memcpy(tmp, rmod + (int64_t)(j - 1) * ni, sizeof(double) * ni);
    lu_solve(w->lu + (int64_t)(j - 1) * ni * ni, ni, w->piv + (int64_t)(j - 1) * ni, tmp);
    for (int i = 0; i < ni; ++i)
      rmod[(int64_t)j * ni + i] = rhs[(int64_t)j * ni + i] - sub_neg[(int64_t)(j - 1) * ni + i] * tmp[i];
  }
  memcpy(x + (int64_t)(nz - 1) * ni, rmod + (int64_t)(nz - 1) * ni, sizeof(double) * ni);
  lu_solve(w->lu + (int64_t)(nz - 1) * ni * ni, ni, w->piv + (int64_t)(nz - 1) * ni,
           x + (int64_t)(nz - 1) * ni);
  for (int j = nz - 2; j >= 0; --j) {
    for (int i = 0; i < ni; ++i)
      x[(int64_t)j * ni + i] = rmod[(int64_t)j * ni + i] - sup_neg[(int64_t)j * ni + i] * x[(int64_t)(j + 1) * ni + i];
    lu_solve(w->lu + (int64_t)j * ni * ni, ni, w->piv + (int64_t)j * ni, x + (int64_t)j * ni);
  }
}

static void ros2_one(const Net* n, const Atm* v, const double* y, const double* k, double dt,
                     double* sol, double* delta, Work* w) {
  const int nz = n->nz, ni = n->ni, W = ni + 1;
  const double r = 1.0 + 1.0 / sqrt(2.0);
  const double c0 = 1.0 / (r * dt);
  total_rhs(n, v, y, k, w->f, w->dd, w->sup, w->sub, w->rate, w->loss, w->prod);
  for (int64_t e = 0; e < (int64_t)(nz - 1) * ni; ++e) { w->sup[e] = -w->sup[e]; w->sub[e] = -w->sub[e]; }
  /* factor: A'_0 = LU(D_0); A'_j = LU(D_j - (c b^T) o inv(A'_{j-1})) */
  for (int z = 0; z < nz; ++z) {
    layer_rates(n, y + (int64_t)z * ni, v->M[z], k, z, w->rate, w->dr);
    jac_layer(n, w->dr, w->J);
    double* A = w->lu + (int64_t)z * ni * ni;
    for (int i = 0; i < ni; ++i)
      for (int q = 0; q < ni; ++q) {
        double d = (i == q ? c0 : 0.0) - w->J[i * W + q];
        if (i == q) d = d + (-w->dd[(int64_t)z * ni + i]);
        A[(int64_t)i * ni + q] = d;
      }
    if (z > 0) {
      lu_inverse(w->lu + (int64_t)(z - 1) * ni * ni, ni, w->piv + (int64_t)(z - 1) * ni, w->inv, w->col);
      const double* c = w->sub + (int64_t)(z - 1) * ni;
      const double* b = w->sup + (int64_t)(z - 1) * ni;
      for (int i = 0; i < ni; ++i)
        for (int q = 0; q < ni; ++q) A[(int64_t)i * ni + q] -= (c[i] * b[q]) * w->inv[(int64_t)i * ni + q];
    }
    lu_factor(A, ni, w->piv + (int64_t)z * ni);
  }
  bt_solve(w, nz, ni, w->sup, w->sub, w->f, w->k1, w->rmod, w->col);
  for (int64_t e = 0; e < (int64_t)nz * ni; ++e) w->yk2[e] = y[e] + w->k1[e] / r;
  total_rhs(n, v, w->yk2, k, w->rhs2, NULL, NULL, NULL, w->rate, w->loss, w->prod);
  const double c2 = 2.0 / (r * dt);
  for (int64_t e = 0; e < (int64_t)nz * ni; ++e) w->rhs2[e] = w->rhs2[e] - c2 * w->k1[e];
  bt_solve(w, nz, ni, w->sup, w->sub, w->rhs2, w->k2, w->rmod, w->col);
  const double e0 = 3.0 / (2.0 * r), e1 = 1.0 / (2.0 * r);
  for (int64_t e = 0; e < (int64_t)nz * ni; ++e) {
    const double s = y[e] + e0 * w->k1[e] + e1 * w->k2[e];
    sol[e] = s;
    delta[e] = fabs(s - w->yk2[e]);
  }
}

/* One Ros2 step for each of ncols columns sharing one network; arrays are
 * stacked per column.  nthreads > 1 runs one column per OpenMP thread. */
int oracle_ros2_steps(int ncols, int nz, int ni, int nr, int T, const int64_t* r_idx,
                      const int64_t* p_idx, const double* r_st, const double* p_st,
                      const uint8_t* is3, const double* y, const double* k, const double* atm,
                      int64_t atm_len, const double* dt, double* sol, double* delta,
                      int nthreads) {
  Net n = {nz, ni, nr + 1, T, r_idx, p_idx, r_st, p_st, is3};
  int err = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1) reduction(| : err)
  {
    Work w;
    if (work_alloc(&w, nz, ni, n.S, T) != 0) err |= 1;
#pragma omp for schedule(dynamic, 1)
    for (int c = 0; c < ncols; ++c) {
      Atm v = atm_view(atm + (int64_t)c * atm_len, nz, ni);
      const int64_t os = (int64_t)c * nz * ni;
      ros2_one(&n, &v, y + os, k + (int64_t)c * n.S * nz, dt[c], sol + os, delta + os, &w);
    }
    work_free(&w);
  }
  return err;
}
