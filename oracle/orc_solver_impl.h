/* Precision template for the oracle operator and restarted GMRES; included by
 * orc_gmres.c with R, SFX defined.  Vectors in R, inner products accumulated
 * in FP64, Hessenberg/Givens/least squares always FP64. */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define CX CAT(cx, SFX)

/* y = x - G(chi x) (mode 0), x - chi G(x) (mode 1), G(x) (mode 2), via the
 * circulant embedding: pad to M*M, 2-D FFT, multiply by Ghat, inverse FFT,
 * crop to N*N, divide by M^2.  B/CUDA_GMRES/Fast_matvec.cuh:1-56 with
 * extend_by_zeros_kernel (GMRES_kernels.cuh:98-149), MatMul_ElemWise_kernel
 * (151-167) and _2D_to_1D_kernel (286-326). */
void FN(orc_matvec_)(const orc_op *op, const R *chi, const R *x_, R *y_, int mode) {
  const int N = op->N, M = op->M;
  const CX *x = (const CX *)x_;
  CX *y = (CX *)y_;
  CX *E = (CX *)calloc((size_t)M * M, sizeof(CX));
  CX *work = (CX *)malloc(sizeof(CX) * M);
  CX *col = (CX *)malloc(sizeof(CX) * M);
  const double *twr = op->tw, *twi = op->tw + M;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const CX v = x[(long)i * N + j];
      const R s = (mode == 0) ? chi[(long)i * N + j] : (R)1;
      E[(long)i * M + j].re = s * v.re;
      E[(long)i * M + j].im = s * v.im;
    }
  /* forward: rows 0..N-1 are the only non-zero rows */
  for (int i = 0; i < N; ++i) FN(fft1)(M, E + (long)i * M, work, -1, twr, twi);
  for (int j = 0; j < M; ++j) {
    for (int i = 0; i < M; ++i) col[i] = E[(long)i * M + j];
    FN(fft1)(M, col, work, -1, twr, twi);
    for (int i = 0; i < M; ++i) {
      const long q = (long)i * M + j;
      CX a = col[i], g;
      if (sizeof(R) == sizeof(float)) { g.re = (R)op->ghat32[2 * q]; g.im = (R)op->ghat32[2 * q + 1]; }
      else { g.re = (R)op->ghat[2 * q]; g.im = (R)op->ghat[2 * q + 1]; }
      col[i].re = a.re * g.re - a.im * g.im;
      col[i].im = a.re * g.im + a.im * g.re;
    }
    FN(fft1)(M, col, work, +1, twr, twi);
    for (int i = 0; i < N; ++i) E[(long)i * M + j] = col[i];
  }
  const R inv = (R)(1.0 / ((double)M * (double)M));
  for (int i = 0; i < N; ++i) {
    FN(fft1)(M, E + (long)i * M, work, +1, twr, twi);
    for (int j = 0; j < N; ++j) {
      const long p = (long)i * N + j;
      const CX t = E[(long)i * M + j];
      CX o;
      if (mode == 2) { o.re = t.re * inv; o.im = t.im * inv; }
      else {
        const R s = (mode == 1) ? chi[p] * inv : inv;
        o.re = x[p].re - s * t.re;
        o.im = x[p].im - s * t.im;
      }
      y[p] = o;
    }
  }
  free(col);
  free(work);
  free(E);
}

static double FN(nrm2)(long n, const CX *v) {
  double s = 0;
  for (long i = 0; i < n; ++i) s += (double)v[i].re * v[i].re + (double)v[i].im * v[i].im;
  return sqrt(s);
}

/* h = sum conj(a) b, FP64 accumulation */
static void FN(dotc)(long n, const CX *a, const CX *b, double *hr, double *hi) {
  double sr = 0, si = 0;
  for (long i = 0; i < n; ++i) {
    sr += (double)a[i].re * b[i].re + (double)a[i].im * b[i].im;
    si += (double)a[i].re * b[i].im - (double)a[i].im * b[i].re;
  }
  *hr = sr; *hi = si;
}

static void FN(apply)(const orc_op *op, const R *chi, const CX *x, CX *y, int adjoint) {
  FN(orc_matvec_)(op, chi, (const R *)x, (R *)y, adjoint ? 1 : 0);
}

int FN(orc_gmres_)(const orc_op *op, const R *chi, const R *b_, R *x_,
                   const orc_gmres_opts *o, orc_gmres_info *info, double *res_hist) {
  const long n = (long)op->N * op->N;
  const int m = o->restart;
  const CX *b = (const CX *)b_;
  CX *x = (CX *)x_;
  CX *V = (CX *)malloc(sizeof(CX) * n * (m + 1));
  CX *w = (CX *)malloc(sizeof(CX) * n);
  double complex *H = (double complex *)calloc((size_t)(m + 1) * m, sizeof(double complex));
  double *cs = (double *)malloc(sizeof(double) * m);
  double complex *sn = (double complex *)malloc(sizeof(double complex) * m);
  double complex *g = (double complex *)malloc(sizeof(double complex) * (m + 1));
  double complex *yv = (double complex *)malloc(sizeof(double complex) * m);
  const double bnorm = FN(nrm2)(n, b);
  int total = 0, hist = 0, cyc = 0;
  info->converged = 0;
  if (bnorm == 0.0) {
    for (long i = 0; i < n; ++i) { x[i].re = 0; x[i].im = 0; }
    info->iterations = 0; info->cycles = 0; info->converged = 1;
    info->rel_residual = 0.0; info->beta0 = 0.0;
    goto done;
  }
  for (cyc = 0; cyc <= o->max_cycles; ++cyc) {
    /* true residual r = b - A x */
    FN(apply)(op, chi, x, w, o->adjoint);
    for (long i = 0; i < n; ++i) { w[i].re = b[i].re - w[i].re; w[i].im = b[i].im - w[i].im; }
    const double beta = FN(nrm2)(n, w);
    if (cyc == 0) info->beta0 = beta;
    info->rel_residual = beta / bnorm;
    if (res_hist && cyc == 0) res_hist[hist++] = beta;
    if (beta / bnorm <= o->tol) { info->converged = 1; break; }
    if (cyc == o->max_cycles) break;
    for (long i = 0; i < n; ++i) { V[i].re = (R)(w[i].re / beta); V[i].im = (R)(w[i].im / beta); }
    for (int i = 0; i <= m; ++i) g[i] = 0;
    g[0] = beta;
    int k = 0;
    for (int i = 0; i < m; ++i) {
      CX *vi = V + (long)i * n, *vn = V + (long)(i + 1) * n;
      FN(apply)(op, chi, vi, vn, o->adjoint);
      /* CGS2: two classical Gram-Schmidt passes */
      for (int j = 0; j <= i; ++j) H[(long)j * m + i] = 0;
      for (int pass = 0; pass < 2; ++pass) {
        double complex hh[512];
        for (int j = 0; j <= i; ++j) {
          double hr, hi;
          FN(dotc)(n, V + (long)j * n, vn, &hr, &hi);
          hh[j] = hr + I * hi;
        }
        for (long e = 0; e < n; ++e) {
          double ar = vn[e].re, ai = vn[e].im;
          for (int j = 0; j <= i; ++j) {
            const CX vj = V[(long)j * n + e];
            ar -= creal(hh[j]) * vj.re - cimag(hh[j]) * vj.im;
            ai -= creal(hh[j]) * vj.im + cimag(hh[j]) * vj.re;
          }
          vn[e].re = (R)ar; vn[e].im = (R)ai;
        }
        for (int j = 0; j <= i; ++j) H[(long)j * m + i] += hh[j];
      }
      const double hn = FN(nrm2)(n, vn);
      H[(long)(i + 1) * m + i] = hn;
      if (hn > 0) {
        const double inv = 1.0 / hn;
        for (long e = 0; e < n; ++e) { vn[e].re = (R)(vn[e].re * inv); vn[e].im = (R)(vn[e].im * inv); }
      }
      /* apply previous rotations to column i */
      for (int j = 0; j < i; ++j) {
        const double complex a = H[(long)j * m + i], c = H[(long)(j + 1) * m + i];
        H[(long)j * m + i] = cs[j] * a + sn[j] * c;
        H[(long)(j + 1) * m + i] = -conj(sn[j]) * a + cs[j] * c;
      }
      /* new unitary rotation [[c, s], [-conj(s), c]], c real */
      {
        const double complex a = H[(long)i * m + i], c = H[(long)(i + 1) * m + i];
        const double aa = cabs(a), cc = cabs(c);
        const double d = hypot(aa, cc);
        if (aa == 0.0) { cs[i] = 0.0; sn[i] = 1.0; H[(long)i * m + i] = c; }
        else {
          cs[i] = aa / d;
          sn[i] = (a / aa) * conj(c) / d;
          H[(long)i * m + i] = (a / aa) * d;
        }
        H[(long)(i + 1) * m + i] = 0;
        g[i + 1] = -conj(sn[i]) * g[i];
        g[i] = cs[i] * g[i];
      }
      ++total;
      k = i + 1;
      const double est = cabs(g[i + 1]);
      if (res_hist) res_hist[hist++] = est;
      if (est / bnorm <= o->tol || hn == 0.0) break;
    }
    /* back substitution R y = g */
    for (int i = k - 1; i >= 0; --i) {
      double complex s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[(long)i * m + j] * yv[j];
      yv[i] = s / H[(long)i * m + i];
    }
    for (long e = 0; e < n; ++e) {
      double ar = x[e].re, ai = x[e].im;
      for (int j = 0; j < k; ++j) {
        const CX vj = V[(long)j * n + e];
        ar += creal(yv[j]) * vj.re - cimag(yv[j]) * vj.im;
        ai += creal(yv[j]) * vj.im + cimag(yv[j]) * vj.re;
      }
      x[e].re = (R)ar; x[e].im = (R)ai;
    }
  }
  info->iterations = total;
  info->cycles = cyc;
done:
  free(yv); free(g); free(sn); free(cs); free(H); free(w); free(V);
  return info->converged ? 0 : 4;
}

#undef FN
#undef CX
#undef CAT
#undef CAT2
