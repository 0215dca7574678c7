/*
 * Problem generators for the oracle (test-only): Green's blocks in reference
 * and baseline modes, circulant embedding, masks and plane waves.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define ORC_PI 3.14159265358979323846

/* B/CUDA_GMRES/GMRES_kernels.cuh:12  WAVE_NUMBER 2*3.14f/(N/6.f), float. */
float orc_ref_wave_number(int N) {
  volatile float two_pi_approx = 2.0f * 3.14f;
  volatile float lam = (float)N / 6.0f;
  return two_pi_approx / lam;
}

/* B/MKL_GMRES_test_09122019/source/loop_functions.h:134-143:
 * kr = WAVE_NUMBER * sqrt(pow((float)i - 0.5f, 2) + pow((float)j, 2)) in float,
 * G = (-0.25 y0(kr), 0.25 j0(kr)) evaluated in double and stored as float. */
void orc_ref_green_block_f32(int N, float *blk) {
  const float k = orc_ref_wave_number(N);
  for (int i = 0; i < N; ++i) {
    for (int j = 0; j < N; ++j) {
      volatile float di = (float)i - 0.5f;
      volatile float dj = (float)j;
      volatile float s = di * di + dj * dj;
      volatile float r = sqrtf(s);
      volatile float kr = k * r;
      double j0, j1, y0, y1;
      orc_bessel01((double)kr, &j0, &j1, &y0, &y1);
      blk[2 * ((long)i * N + j)] = (float)(-0.25 * y0);
      blk[2 * ((long)i * N + j) + 1] = (float)(0.25 * j0);
    }
  }
}

/* loop_functions.h:150-163: mirror (i,j) into [M'-i], [M'-j] of the
 * (2N-1)^2 array, M' = 2N-1. */
void orc_ref_gprep_f32(int N, float *gprep) {
  const int Mp = 2 * N - 1;
  float *blk = (float *)malloc(sizeof(float) * 2 * (size_t)N * N);
  orc_ref_green_block_f32(N, blk);
  memset(gprep, 0, sizeof(float) * 2 * (size_t)Mp * Mp);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const float re = blk[2 * ((long)i * N + j)], im = blk[2 * ((long)i * N + j) + 1];
      long idx[4];
      int cnt = 0;
      idx[cnt++] = (long)i * Mp + j;
      if (i > 0 && j > 0) idx[cnt++] = (long)(Mp - i) * Mp + (Mp - j);
      if (i > 0) idx[cnt++] = (long)(Mp - i) * Mp + j;
      if (j > 0) idx[cnt++] = (long)i * Mp + (Mp - j);
      for (int c = 0; c < cnt; ++c) { gprep[2 * idx[c]] = re; gprep[2 * idx[c] + 1] = im; }
    }
  free(blk);
}

/* k0^2 G for the h^2-weighted Hankel operator of the benchmark, FP64.  rho in cells, kh = k0 h. */
void orc_baseline_green_block(int N, double kh, double *blk) {
  const double ka = kh / sqrt(ORC_PI);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double re, im;
      if (i == 0 && j == 0) {
        double j0, j1, y0, y1;
        orc_bessel01(ka, &j0, &j1, &y0, &y1);
        /* (i pi ka / 2) (J1 + i Y1) - 1 */
        re = -0.5 * ORC_PI * ka * y1 - 1.0;
        im = 0.5 * ORC_PI * ka * j1;
      } else {
        double j0, j1, y0, y1;
        const double rho = sqrt((double)i * i + (double)j * j);
        orc_bessel01(kh * rho, &j0, &j1, &y0, &y1);
        /* kh^2 (i/4)(J0 + i Y0) */
        re = -0.25 * kh * kh * y0;
        im = 0.25 * kh * kh * j0;
      }
      blk[2 * ((long)i * N + j)] = re;
      blk[2 * ((long)i * N + j) + 1] = im;
    }
}

void orc_shifted_green_block(int N, double k, double sr, double sc, double *blk) {
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double di = i - sr, dj = j - sc;
      double j0, j1, y0, y1;
      orc_bessel01(k * sqrt(di * di + dj * dj), &j0, &j1, &y0, &y1);
      blk[2 * ((long)i * N + j)] = -0.25 * y0;
      blk[2 * ((long)i * N + j) + 1] = 0.25 * j0;
    }
}

int orc_gprep_to_block(int N, const float *gprep, double *blk) {
  const int Mp = 2 * N - 1;
  int sym_ok = 1;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const long a = (long)i * Mp + j;
      blk[2 * ((long)i * N + j)] = gprep[2 * a];
      blk[2 * ((long)i * N + j) + 1] = gprep[2 * a + 1];
      if (i > 0) {
        const long b = (long)(Mp - i) * Mp + j;
        if (gprep[2 * b] != gprep[2 * a] || gprep[2 * b + 1] != gprep[2 * a + 1]) sym_ok = 0;
      }
      if (j > 0) {
        const long b = (long)i * Mp + (Mp - j);
        if (gprep[2 * b] != gprep[2 * a] || gprep[2 * b + 1] != gprep[2 * a + 1]) sym_ok = 0;
      }
      if (i > 0 && j > 0) {
        const long b = (long)(Mp - i) * Mp + (Mp - j);
        if (gprep[2 * b] != gprep[2 * a] || gprep[2 * b + 1] != gprep[2 * a + 1]) sym_ok = 0;
      }
    }
  return sym_ok ? 0 : 1;
}

void orc_embed_block(int N, int M, const double *blk, double *E) {
  memset(E, 0, sizeof(double) * 2 * (size_t)M * M);
  for (int di = -(N - 1); di <= N - 1; ++di)
    for (int dj = -(N - 1); dj <= N - 1; ++dj) {
      const int ai = di < 0 ? -di : di, aj = dj < 0 ? -dj : dj;
      const long src = (long)ai * N + aj;
      const long dst = (long)((di + M) % M) * M + ((dj + M) % M);
      E[2 * dst] = blk[2 * src];
      E[2 * dst + 1] = blk[2 * src + 1];
    }
}

/* B/PyGMRES/GMRES.py:104-106 (transposed into the C++ row/col orientation,
 * which is the orientation of B/CUDA_GMRES/Release/input/cylinder_N.txt):
 * (i - N/3)^2 + (j - N/2)^2 <= (N/6)^2 in FP64. */
void orc_ref_cylinder_mask(int N, unsigned char *mask) {
  const double c0 = N / 3.0, c1 = N / 2.0, r = N / 6.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double a = i - c0, b = j - c1;
      mask[(long)i * N + j] = (a * a + b * b <= r * r) ? 1 : 0;
    }
}

/* B/CUDA_GMRES/GMRES_kernels.cuh:557-572 init_x0_kernel with ALPHA = 0:
 * phase = (float)(-(double)WAVE_NUMBER * i), then sincosf. */
void orc_ref_plane_wave_f32(int N, float *b) {
  const float k = orc_ref_wave_number(N);
  for (int i = 0; i < N; ++i) {
    const float ph = (float)(-(double)k * ((double)i * 1.0 + 0.0));
    const float c = cosf(ph), s = sinf(ph);
    for (int j = 0; j < N; ++j) {
      b[2 * ((long)i * N + j)] = c;
      b[2 * ((long)i * N + j) + 1] = s;
    }
  }
}

void orc_plane_wave(int N, double kh, double theta, double *b) {
  const double ct = cos(theta), st = sin(theta);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const double ph = -kh * (i * ct + j * st);
      b[2 * ((long)i * N + j)] = cos(ph);
      b[2 * ((long)i * N + j) + 1] = sin(ph);
    }
}
