// nm_soft.hip -- differentiable expected-depth ray march
// (include/nanomap_amd_soft.h).
//
// One lane per ray, 16x16 pixel tiles per workgroup (8x8 per wavefront),
// cameras along the grid's y dimension, so a whole batch of frames is one
// launch.  Per sample the log-odds come from a quadratic B-spline over the
// dense grid (27 taps, value and gradient from the same taps).
//
// The adjoint re-marches each ray from its far end.  With T_k the
// transmittance before sample k and E_k = p_k d_k + (1 - p_k) E_{k+1} the
// expected depth of the samples from k on, given the ray reaches k,
//   D = sum_{i<k} w_i d_i + T_k E_k,  so  dD/dp_k = T_k (d_k - E_{k+1}),
// and dp/dl = p (1 - p).  E is built by the backward recursion (positive
// terms, no cancellation); T_k = T_N / prod_{j>=k} (1 - p_j), from the
// total log-transmittance log T_N that the forward pass stores per ray (one
// number per ray, no trace).  The FP32 kernels carry transmittances as a
// mantissa and a binary exponent (no underflow over any number of samples)
// and take p and 1 - p from one exponential per sample; the FP64 kernels
// sum log(1 - p) = -softplus(l) directly.
#include "../../include/nanomap_amd_soft.h"
#include "nm_internal.h"

namespace nm {
namespace {

template <typename T>
struct SoftArgs {
  const T* grid;
  int nx, ny, nz;
  double ox, oy, oz;  // grid origin (m)
  double inv_res;
  T res;
  T bg;
  const nm_soft_camera* cams;
  int width, height;
  T tx, ty;  // tan of the half fields of view
  T step;
  int ns;
};

template <typename T>
__device__ __forceinline__ T sigmoid_(T l);
template <>
__device__ __forceinline__ float sigmoid_(float l) {
  return 1.0f / (1.0f + expf(-l));
}
template <>
__device__ __forceinline__ double sigmoid_(double l) {
  return 1.0 / (1.0 + exp(-l));
}

template <typename T>
__device__ __forceinline__ T exp_t(T x);
template <>
__device__ __forceinline__ float exp_t(float x) {
  return expf(x);
}
template <>
__device__ __forceinline__ double exp_t(double x) {
  return exp(x);
}

template <typename T>
__device__ __forceinline__ T floor_t(T x);
template <>
__device__ __forceinline__ float floor_t(float x) {
  return __builtin_floorf(x);
}
template <>
__device__ __forceinline__ double floor_t(double x) {
  return __builtin_floor(x);
}

template <typename T>
__device__ __forceinline__ T sqrt_t(T x);
template <>
__device__ __forceinline__ float sqrt_t(float x) {
  return __builtin_sqrtf(x);
}
template <>
__device__ __forceinline__ double sqrt_t(double x) {
  return __builtin_sqrt(x);
}

// Quadratic B-spline weights (and derivatives) of the three taps around
// the voxel holding coordinate q: offset s = q - floor(q) - 1/2.
template <typename T>
__device__ __forceinline__ void bspline2(T s, T w[3], T dw[3]) {
  const T a = T(0.5) - s, b = T(0.5) + s;
  w[0] = T(0.5) * a * a;
  w[1] = T(0.75) - s * s;
  w[2] = T(0.5) * b * b;
  dw[0] = -a;
  dw[1] = T(-2) * s;
  dw[2] = b;
}

// The spline's taps at grid coordinates q (voxel units).
template <typename T>
struct Taps {
  int b[3];  // voxel holding q; taps at b - 1 .. b + 1 per axis
  T w[3][3], dw[3][3];
};

template <typename T>
__device__ __forceinline__ void taps_at(const T q[3], Taps<T>& tp) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T f = floor_t(q[a]);
    tp.b[a] = (int)f;
    bspline2(q[a] - f - T(0.5), tp.w[a], tp.dw[a]);
  }
}

template <typename T>
__device__ __forceinline__ T coef(const SoftArgs<T>& A, int i, int j, int k,
                                  bool inside) {
  if (inside) return A.grid[((size_t)i * A.ny + j) * A.nz + k];
  if ((unsigned)i >= (unsigned)A.nx || (unsigned)j >= (unsigned)A.ny ||
      (unsigned)k >= (unsigned)A.nz)
    return A.bg;
  return A.grid[((size_t)i * A.ny + j) * A.nz + k];
}

template <typename T>
__device__ __forceinline__ bool taps_inside(const SoftArgs<T>& A,
                                            const Taps<T>& tp) {
  return tp.b[0] >= 1 && tp.b[0] + 1 < A.nx && tp.b[1] >= 1 &&
         tp.b[1] + 1 < A.ny && tp.b[2] >= 1 && tp.b[2] + 1 < A.nz;
}

// Three consecutive taps along z (k - 1 .. k + 1) of row (i, j).  Inside
// the grid an FP32 row is one 12-byte load.
typedef float float3_u __attribute__((ext_vector_type(3), aligned(4)));

template <typename T>
__device__ __forceinline__ void row3(const SoftArgs<T>& A, int i, int j,
                                     int k, bool inside, T C[3]) {
  if constexpr (sizeof(T) == 4) {
    if (inside) {
      const float3_u r = *reinterpret_cast<const float3_u*>(
          A.grid + ((size_t)i * A.ny + j) * A.nz + (k - 1));
      C[0] = r.x;
      C[1] = r.y;
      C[2] = r.z;
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) C[c] = coef(A, i, j, k + c - 1, inside);
}

// Value and gradient (per voxel) of the spline.
template <typename T>
__device__ __forceinline__ T spline(const SoftArgs<T>& A, const Taps<T>& tp,
                                    T g[3]) {
  const bool in = taps_inside(A, tp);
  T v = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T r = 0, rz = 0;
      T C3[3];
      row3(A, tp.b[0] + a - 1, tp.b[1] + b - 1, tp.b[2], in, C3);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T C = C3[c];
        r += tp.w[2][c] * C;
        rz += tp.dw[2][c] * C;
      }
      v += tp.w[0][a] * tp.w[1][b] * r;
      gx += tp.dw[0][a] * tp.w[1][b] * r;
      gy += tp.w[0][a] * tp.dw[1][b] * r;
      gz += tp.w[0][a] * tp.w[1][b] * rz;
    }
  }
  g[0] = gx;
  g[1] = gy;
  g[2] = gz;
  return v;
}

// Value only (the forward pass).
template <typename T>
__device__ __forceinline__ T spline_value(const SoftArgs<T>& A,
                                          const Taps<T>& tp) {
  const bool in = taps_inside(A, tp);
  T v = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      T r = 0;
      T C3[3];
      row3(A, tp.b[0] + a - 1, tp.b[1] + b - 1, tp.b[2], in, C3);
#pragma unroll
      for (int c = 0; c < 3; ++c) r += tp.w[2][c] * C3[c];
      v += tp.w[0][a] * tp.w[1][b] * r;
    }
  }
  return v;
}

// Ray of thread `tid` in tile `t` of camera `c`: eye and direction in grid
// voxel units (e, dv) and the world direction (dir).  False outside the
// image.
template <typename T>
__device__ __forceinline__ bool ray_setup(const SoftArgs<T>& A, int c, int t,
                                          int tid, T e[3], T dv[3], T dir[3],
                                          int* pix) {
  const int tiles_x = (A.width + 15) >> 4;
  const int w = tid >> 6, lane = tid & 63;
  const int u = (t % tiles_x) * 16 + (w & 1) * 8 + (lane & 7);
  const int v = (t / tiles_x) * 16 + (w >> 1) * 8 + (lane >> 3);
  if (u >= A.width || v >= A.height) return false;
  *pix = v * A.width + u;
  const nm_soft_camera& cam = A.cams[c];
  const T xc = (T(2) * (T(u) + T(0.5)) / T(A.width) - T(1)) * A.tx;
  const T yc = (T(2) * (T(v) + T(0.5)) / T(A.height) - T(1)) * A.ty;
  const T inv = T(1) / sqrt_t(xc * xc + yc * yc + T(1));
  const T lx = xc * inv, ly = yc * inv, lz = inv;
  const double o[3] = {A.ox, A.oy, A.oz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dir[a] = T(cam.rotation[3 * a + 0]) * lx + T(cam.rotation[3 * a + 1]) * ly +
             T(cam.rotation[3 * a + 2]) * lz;
    e[a] = T((cam.position[a] - o[a]) * A.inv_res);
    dv[a] = dir[a] * T(A.inv_res);
  }
  return true;
}

// log(1 - sigmoid(l)) = -softplus(l), finite for every finite l.
template <typename T>
__device__ __forceinline__ T log1m_sigmoid(T l);
template <>
__device__ __forceinline__ float log1m_sigmoid(float l) {
  return l > 0.0f ? -(l + log1pf(expf(-l))) : -log1pf(expf(l));
}
template <>
__device__ __forceinline__ double log1m_sigmoid(double l) {
  return l > 0.0 ? -(l + log1p(exp(-l))) : -log1p(exp(l));
}

// FP32 occupancy of a sample from one exponential: p = sigmoid(l) and
// q = 1 - p (each without cancellation), z = exp(-min(|l|, 80)).  The
// clamp keeps q a normal float, so transmittances can be carried as a
// mantissa and a binary exponent; it moves p and q by less than e^-80.
__device__ __forceinline__ void occ32(float l, float& p, float& q, float& z,
                                      bool& pos) {
  pos = l >= 0.0f;
  z = __expf(-fminf(fabsf(l), 80.0f));
  const float inv = __builtin_amdgcn_rcpf(1.0f + z);
  p = (pos ? 1.0f : z) * inv;
  q = (pos ? z : 1.0f) * inv;
}

// m <- m * r renormalized to [0.5, 1), with the binary exponent in e.
__device__ __forceinline__ void renorm_mul(float& m, int& e, float r) {
  const float v = m * r;
  e += __builtin_amdgcn_frexp_expf(v);
  m = __builtin_amdgcn_frexp_mantf(v);
}

template <typename T>
__global__ __launch_bounds__(256) void k_soft_forward(SoftArgs<T> A,
                                                      T* __restrict__ depth,
                                                      double* __restrict__ logt) {
  const int c = blockIdx.y;
  T e[3], dv[3], dir[3];
  int pix;
  if (!ray_setup(A, c, blockIdx.x, threadIdx.x, e, dv, dir, &pix)) return;
  T tr = 1, D = 0;
  double lt = 0.0;            // FP64: sum of log(1 - p)
  float tm = 1.0f;            // FP32: transmittance as tm * 2^te
  int te = 0;
  for (int i = 0; i < A.ns; ++i) {
    const T d = (T(i) + T(0.5)) * A.step;
    const T q[3] = {e[0] + dv[0] * d, e[1] + dv[1] * d, e[2] + dv[2] * d};
    Taps<T> tp;
    taps_at(q, tp);
    const T l = spline_value(A, tp);
    if constexpr (sizeof(T) == 4) {
      float p, om, z;
      bool pos;
      occ32(l, p, om, z, pos);
      D += p * tr * d;
      tr *= om;
      renorm_mul(tm, te, om);
    } else {
      const T p = sigmoid_(l);
      const T w = p * tr;
      D += w * d;
      tr *= T(1) - p;
      if (logt) lt += (double)log1m_sigmoid(l);
    }
  }
  const size_t ray = (size_t)c * A.width * A.height + pix;
  depth[ray] = D;
  if constexpr (sizeof(T) == 4)
    lt = (double)logf(tm) + (double)te * 0.69314718055994531;
  if (logt) logt[ray] = lt;
}

__device__ __forceinline__ void atomic_add_t(float* p, float v) {
  atomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_t(double* p, double v) {
  atomicAdd(p, v);
}

template <typename T, typename G, bool GRID>
__global__ __launch_bounds__(256) void k_soft_backward(
    SoftArgs<T> A, const double* __restrict__ logt,
    const T* __restrict__ gdepth, double* __restrict__ gpose,
    G* __restrict__ gray, T* __restrict__ ggrid) {
  __shared__ double s_pose[4][6];
  const int c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T e[3], dv[3], dir[3];
  int pix = 0;
  const bool valid = ray_setup(A, c, blockIdx.x, threadIdx.x, e, dv, dir, &pix);
  T jr[3] = {0, 0, 0}, jw[3] = {0, 0, 0};  // dD/drho, dD/domega
  T gout = 0;
  if (valid) {
    const size_t ray = (size_t)c * A.width * A.height + pix;
    double lt = logt[ray];  // log T_N, then log T_k on the way back
    // FP32: T_k as tm * 2^te, from T_N up by T_k = T_{k+1} / (1 - p_k)
    const double y = lt * 1.4426950408889634;  // log2 T_N
    const double fy = floor(y);
    float tm = exp2f((float)(y - fy));
    int te = (int)fy;
    gout = gdepth ? gdepth[ray] : T(0);
    const T inv_res = T(A.inv_res);
    T E = 0;  // E_{k+1}
    for (int k = A.ns - 1; k >= 0; --k) {
      const T d = (T(k) + T(0.5)) * A.step;
      const T q[3] = {e[0] + dv[0] * d, e[1] + dv[1] * d, e[2] + dv[2] * d};
      Taps<T> tp;
      taps_at(q, tp);
      T g[3];
      const T l = spline(A, tp, g);
      T p, om, tk;
      if constexpr (sizeof(T) == 4) {
        float z;
        bool pos;
        occ32(l, p, om, z, pos);
        // 1 / (1 - p) = 1 + e^l
        renorm_mul(tm, te, pos ? 1.0f + __builtin_amdgcn_rcpf(z) : 1.0f + z);
        tk = __builtin_amdgcn_ldexpf(tm, te);
      } else {
        p = sigmoid_(l);
        om = T(1) - p;
        lt -= (double)log1m_sigmoid(l);
        tk = exp_t((T)lt);
      }
      const T dDdl = tk * (d - E) * p * om;
      E = p * d + om * E;
      if (dDdl == T(0)) continue;
      // world gradient of l and the sample's offset from the eye
      const T gx = g[0] * inv_res, gy = g[1] * inv_res, gz = g[2] * inv_res;
      const T rx = dir[0] * d, ry = dir[1] * d, rz = dir[2] * d;
      jr[0] += dDdl * gx;
      jr[1] += dDdl * gy;
      jr[2] += dDdl * gz;
      jw[0] += dDdl * (ry * gz - rz * gy);
      jw[1] += dDdl * (rz * gx - rx * gz);
      jw[2] += dDdl * (rx * gy - ry * gx);
      if (GRID && gout != T(0)) {
        const T s = gout * dDdl;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
              const int i0 = tp.b[0] + a - 1, j0 = tp.b[1] + b - 1,
                        k0 = tp.b[2] + cc - 1;
              if ((unsigned)i0 < (unsigned)A.nx &&
                  (unsigned)j0 < (unsigned)A.ny &&
                  (unsigned)k0 < (unsigned)A.nz)
                atomic_add_t(&ggrid[((size_t)i0 * A.ny + j0) * A.nz + k0],
                             s * tp.w[0][a] * tp.w[1][b] * tp.w[2][cc]);
            }
      }
    }
    if (gray) {
      G* o = gray + ray * 6;
      o[0] += (G)jr[0];
      o[1] += (G)jr[1];
      o[2] += (G)jr[2];
      o[3] += (G)jw[0];
      o[4] += (G)jw[1];
      o[5] += (G)jw[2];
    }
  }
  if (!gpose) return;
  // camera pose gradient: sum over the workgroup's rays of gout * dD/dxi
  double v[6] = {(double)(gout * jr[0]), (double)(gout * jr[1]),
                 (double)(gout * jr[2]), (double)(gout * jw[0]),
                 (double)(gout * jw[1]), (double)(gout * jw[2])};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
    if (lane == 0) s_pose[wave][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const double s = s_pose[0][threadIdx.x] + s_pose[1][threadIdx.x] +
                     s_pose[2][threadIdx.x] + s_pose[3][threadIdx.x];
    if (s != 0.0) atomicAdd(&gpose[(size_t)c * 6 + threadIdx.x], s);
  }
}

template <typename T>
int make_args(const nm_soft_grid* g, const T* logodds,
              const nm_soft_camera* cams, int n, const nm_soft_params* p,
              SoftArgs<T>* A) {
  if (!g || !p || !cams || !logodds || n <= 0) return NM_EINVAL;
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || !(g->res > 0)) return NM_EINVAL;
  if (p->width <= 0 || p->height <= 0 || p->n_samples <= 0 || !(p->step > 0) ||
      !(p->hfov_deg > 0 && p->hfov_deg < 180) ||
      !(p->vfov_deg > 0 && p->vfov_deg < 180) || n > 65535)
    return NM_EINVAL;
  A->grid = logodds;
  A->nx = g->nx;
  A->ny = g->ny;
  A->nz = g->nz;
  A->ox = g->origin[0];
  A->oy = g->origin[1];
  A->oz = g->origin[2];
  A->inv_res = 1.0 / (double)g->res;
  A->res = (T)g->res;
  A->bg = (T)g->background;
  A->cams = cams;
  A->width = p->width;
  A->height = p->height;
  const double pi = 3.14159265358979323846;
  A->tx = (T)tan(p->hfov_deg * pi / 360.0);
  A->ty = (T)tan(p->vfov_deg * pi / 360.0);
  A->step = (T)p->step;
  A->ns = p->n_samples;
  return NM_OK;
}

dim3 soft_grid(const nm_soft_params* p, int n) {
  return dim3(((p->width + 15) >> 4) * ((p->height + 15) >> 4), n);
}

}  // namespace
}  // namespace nm

extern "C" {

int nm_soft_forward(const nm_soft_grid* g, const float* logodds,
                    const nm_soft_camera* cams, int n,
                    const nm_soft_params* p, float* depth,
                    double* log_transmittance, void* stream) {
  nm::SoftArgs<float> A;
  if (nm::make_args(g, logodds, cams, n, p, &A) != NM_OK || !depth) {
    nm::set_error("nm_soft_forward: bad argument");
    return NM_EINVAL;
  }
  nm::k_soft_forward<float><<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
      A, depth, log_transmittance);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    nm::set_error(std::string("nm_soft_forward: ") + hipGetErrorString(e));
    return NM_EHIP;
  }
  return NM_OK;
}

int nm_soft_backward(const nm_soft_grid* g, const float* logodds,
                     const nm_soft_camera* cams, int n,
                     const nm_soft_params* p, const double* log_transmittance,
                     const float* grad_depth, double* grad_pose,
                     float* grad_ray_pose, float* grad_logodds,
                     void* stream) {
  nm::SoftArgs<float> A;
  if (nm::make_args(g, logodds, cams, n, p, &A) != NM_OK || !log_transmittance ||
      ((grad_pose || grad_logodds) && !grad_depth)) {
    nm::set_error("nm_soft_backward: bad argument");
    return NM_EINVAL;
  }
  if (grad_logodds)
    nm::k_soft_backward<float, float, true>
        <<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
            A, log_transmittance, grad_depth, grad_pose, grad_ray_pose,
            grad_logodds);
  else
    nm::k_soft_backward<float, float, false>
        <<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
            A, log_transmittance, grad_depth, grad_pose, grad_ray_pose,
            nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    nm::set_error(std::string("nm_soft_backward: ") + hipGetErrorString(e));
    return NM_EHIP;
  }
  return NM_OK;
}

int nm_soft_forward_f64(const nm_soft_grid* g, const double* logodds,
                        const nm_soft_camera* cams, int n,
                        const nm_soft_params* p, double* depth,
                        double* log_transmittance, void* stream) {
  nm::SoftArgs<double> A;
  if (nm::make_args(g, logodds, cams, n, p, &A) != NM_OK || !depth) {
    nm::set_error("nm_soft_forward_f64: bad argument");
    return NM_EINVAL;
  }
  nm::k_soft_forward<double><<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
      A, depth, log_transmittance);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    nm::set_error(std::string("nm_soft_forward_f64: ") + hipGetErrorString(e));
    return NM_EHIP;
  }
  return NM_OK;
}

int nm_soft_backward_f64(const nm_soft_grid* g, const double* logodds,
                         const nm_soft_camera* cams, int n,
                         const nm_soft_params* p,
                         const double* log_transmittance,
                         const double* grad_depth, double* grad_pose,
                         double* grad_logodds, void* stream) {
  nm::SoftArgs<double> A;
  if (nm::make_args(g, logodds, cams, n, p, &A) != NM_OK || !log_transmittance ||
      ((grad_pose || grad_logodds) && !grad_depth)) {
    nm::set_error("nm_soft_backward_f64: bad argument");
    return NM_EINVAL;
  }
  if (grad_logodds)
    nm::k_soft_backward<double, double, true>
        <<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
            A, log_transmittance, grad_depth, grad_pose, nullptr,
            grad_logodds);
  else
    nm::k_soft_backward<double, double, false>
        <<<nm::soft_grid(p, n), 256, 0, (hipStream_t)stream>>>(
            A, log_transmittance, grad_depth, grad_pose, nullptr, nullptr);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    nm::set_error(std::string("nm_soft_backward_f64: ") + hipGetErrorString(e));
    return NM_EHIP;
  }
  return NM_OK;
}

}  // extern "C"
