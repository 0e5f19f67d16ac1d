// Step 14 of FS_GLOBAL on the device: the barotropic correction around the
// solve (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1154-1283).
//
// Every element's value is a function of its own inputs only (the W column
// of one cell, the corrections of one face), so one thread per (I, J)
// performs exactly the reference's operation sequence for that element, in
// REAL(4) arithmetic with REAL(8) where Fortran promotes (P0 += ODT*X, the
// gradient of X); the library is built with -ffp-contract=off.  The largest
// top-layer land speed is a max, reduced per workgroup and then by one
// workgroup, so its value does not depend on the order.
#include "sevp_dev.h"
#include "sevp_internal.h"

namespace sevp {

#define U3(a, i, j, k) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0 + ((long)(k) - 1) * (long)I0 * J0]
#define V3(a, i, j, k) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0 + ((long)(k) - 1) * (long)I0 * J1]
#define A2(a, i, j) (a)[((long)(i) - 1) + ((long)(j) - 1) * I0]

constexpr int S14_TX = 64, S14_TY = 4;

// Land masking of the top-layer velocities (:1158-1171).
__global__ void s14_mask_kernel(Step14Args p) {
  const int I0 = p.i2 + 2, J0 = p.j2 + 2, I1 = I0 - 1, J1 = J0 - 1, I2 = p.i2, J2 = p.j2;
  const int I = blockIdx.x * S14_TX + threadIdx.x + 1;
  const int J = blockIdx.y * S14_TY + threadIdx.y + 1;
  if (I > I0 || J > J0) return;
  if (J >= 2 && J <= J1) {
    const bool in = (I >= 2 && I <= I2) || (p.lopenw == 1 && (I == 1 || I == I1));
    if (in) U3(p.u, I, J, 1) = (float)A2(p.iu0, I, J) * U3(p.u, I, J, 1);
  }
  if (J >= 2 && J <= J2 && I >= 2 && I <= I1) V3(p.v, I, J, 1) = (float)A2(p.iv0, I, J) * V3(p.v, I, J, 1);
}

// The shipped source's override U = 1, V = 0, W(:,:,K0) = 0 (:1173-1176).
__global__ void s14_harness_kernel(Step14Args p) {
  const int I0 = p.i2 + 2, J0 = p.j2 + 2, J1 = J0 - 1, K0 = p.k1 + 1;
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nu = (long)I0 * J0 * p.k1, nv = (long)I0 * J1 * p.k1, nw = (long)I0 * J0;
  if (n < nu) p.u[n] = 1.0f;
  if (n < nv) p.v[n] = 0.0f;
  if (n < nw) p.w[n + (long)(K0 - 1) * I0 * J0] = 0.0f;
}

// W integration over the layers of one cell and the RHS S (:1178-1209).
__global__ void s14_rhs_kernel(Step14Args p) {
  const int I0 = p.i2 + 2, J0 = p.j2 + 2, J1 = J0 - 1, I2 = p.i2, K1 = p.k1;
  const int I = blockIdx.x * S14_TX + threadIdx.x + 2;
  const int J = blockIdx.y * S14_TY + threadIdx.y + 2;
  if (I > I0 - 1 || J > J1) return;
  const float odx = p.odx[J - 1], temp = p.ocs[J - 1] * p.ody[J - 1];
  const float csv = p.csv[J - 1], csvm = p.csv[J - 2];
  if (p.free_surface) {
    float wk = U3(p.w, I, J, K1 + 1);
    for (int K = K1; K >= 1; --K) {
      const float temp1 = 1.0f / p.odz[K - 1];
      const float a = (U3(p.u, I, J, K) - U3(p.u, I - 1, J, K)) * odx;
      const float b = (csv * V3(p.v, I, J, K) - csvm * V3(p.v, I, J - 1, K)) * temp;
      wk = wk + (a + b) * temp1;
      U3(p.w, I, J, K) = wk;
    }
    p.s[(I - 2) + (long)(J - 2) * I2] = wk;
  } else {
    float wk = U3(p.w, I, J, 1);
    for (int K = 1; K <= K1; ++K) {
      const float temp1 = 1.0f / p.odz[K - 1];
      const float a = (U3(p.u, I, J, K) - U3(p.u, I - 1, J, K)) * odx;
      const float b = (csv * V3(p.v, I, J, K) - csvm * V3(p.v, I, J - 1, K)) * temp;
      wk = wk - (a + b) * temp1;
      U3(p.w, I, J, K + 1) = wk;
    }
    p.s[(I - 2) + (long)(J - 2) * I2] = -U3(p.w, I, J, A2(p.kb, I, J) + 1);
  }
}

// P0, the U and V corrections with their running sums, and this
// workgroup's largest top-layer land speed (:1214-1283).
__global__ __launch_bounds__(S14_TX *S14_TY) void s14_correct_kernel(Step14Args p) {
  __shared__ float red[S14_TX * S14_TY];
  const int I0 = p.i2 + 2, J0 = p.j2 + 2, I1 = I0 - 1, J1 = J0 - 1, I2 = p.i2, J2 = p.j2;
  const int K1 = p.k1;
  const int I = blockIdx.x * S14_TX + threadIdx.x + 1;
  const int J = blockIdx.y * S14_TY + threadIdx.y + 1;
  float tmax = 0.0f;
  if (I <= I1 && J <= J1) {
    if (I >= 2 && J >= 2) A2(p.p0, I, J) = (float)((double)A2(p.p0, I, J) + (double)p.odt * A2(p.x, I, J));
    if (J >= 2) {  // U faces I = 1..I1 of row J
      float s1 = 0.0f;
      if (I >= 2 && I <= I2)
        s1 = (float)(-((A2(p.x, I + 1, J) - A2(p.x, I, J)) * (double)p.odx[J - 1]));
      else if (p.periodic_we)
        s1 = (float)(-((A2(p.x, 2, J) - A2(p.x, I1, J)) * (double)p.odx[J - 1]));
      U3(p.scr, I, J, 1) = s1;
      U3(p.scr, I, J, 3) = U3(p.scr, I, J, 3) + s1;
      U3(p.u, I, J, 1) = U3(p.u, I, J, 1) + s1;
      for (int K = 2; K <= K1; ++K)
        U3(p.u, I, J, K) = U3(p.u, I, J, K) + s1 * (float)U3(p.iu, I, J, K);
      if (I >= 2 && I <= I2) tmax = (float)(1 - A2(p.iu0, I, J)) * fabsf(U3(p.u, I, J, 1));
    }
    if (I >= 2) {  // V faces J = 1..J1 of column I
      float s2 = 0.0f;
      if (J >= 2 && J <= J2)
        s2 = (float)(-((A2(p.x, I, J + 1) - A2(p.x, I, J)) * (double)p.odyv[J - 1]));
      U3(p.scr, I, J, 2) = s2;
      U3(p.scr, I, J, 4) = U3(p.scr, I, J, 4) + s2;
      V3(p.v, I, J, 1) = V3(p.v, I, J, 1) + s2;
      for (int K = 2; K <= K1; ++K)
        V3(p.v, I, J, K) = V3(p.v, I, J, K) + s2 * (float)V3(p.iv, I, J, K);
      if (J >= 2 && J <= J2) {
        const float t = (float)(1 - A2(p.iv0, I, J)) * fabsf(V3(p.v, I, J, 1));
        tmax = tmax > t ? tmax : t;
      }
    }
  }
  const int t = threadIdx.y * S14_TX + threadIdx.x;
  red[t] = tmax;
  __syncthreads();
  for (int h = S14_TX * S14_TY / 2; h > 0; h >>= 1) {
    if (t < h) red[t] = red[t] > red[t + h] ? red[t] : red[t + h];
    __syncthreads();
  }
  if (t == 0) p.tmp_part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

__global__ void s14_max_kernel(const float *part, long n, float *out) {
  __shared__ float red[256];
  float m = 0.0f;
  for (long i = threadIdx.x; i < n; i += 256) m = m > part[i] ? m : part[i];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + h] ? red[threadIdx.x] : red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void s14_zero_scr_kernel(float *scr, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) scr[i] = 0.0f;
}

#undef U3
#undef V3
#undef A2

long step14_partials(int i2, int j2) {
  const long gx = (i2 + 2 + S14_TX - 1) / S14_TX, gy = (j2 + 2 + S14_TY - 1) / S14_TY;
  return gx * gy;
}

hipError_t launch_step14_prepare(const Step14Args &p, hipStream_t st) {
  const dim3 blk(S14_TX, S14_TY);
  const dim3 g0((p.i2 + 2 + S14_TX - 1) / S14_TX, (p.j2 + 2 + S14_TY - 1) / S14_TY);
  if (p.mask) hipLaunchKernelGGL(s14_mask_kernel, g0, blk, 0, st, p);
  if (p.harness) {
    const long n = (long)(p.i2 + 2) * (p.j2 + 2) * (p.k1 > 1 ? p.k1 : 1);
    hipLaunchKernelGGL(s14_harness_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p);
  }
  const dim3 g1((p.i2 + S14_TX - 1) / S14_TX, (p.j2 + S14_TY - 1) / S14_TY);
  hipLaunchKernelGGL(s14_rhs_kernel, g1, blk, 0, st, p);
  return hipGetLastError();
}

hipError_t launch_step14_correct(const Step14Args &p, float *tmp_dev, hipStream_t st) {
  const dim3 blk(S14_TX, S14_TY);
  const dim3 g0((p.i2 + 2 + S14_TX - 1) / S14_TX, (p.j2 + 2 + S14_TY - 1) / S14_TY);
  hipLaunchKernelGGL(s14_correct_kernel, g0, blk, 0, st, p);
  hipLaunchKernelGGL(s14_max_kernel, dim3(1), dim3(256), 0, st, p.tmp_part,
                     step14_partials(p.i2, p.j2), tmp_dev);
  return hipGetLastError();
}

hipError_t launch_step14_zero_scr(float *scr, long n, hipStream_t st) {
  hipLaunchKernelGGL(s14_zero_scr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, scr, n);
  return hipGetLastError();
}

}  // namespace sevp
