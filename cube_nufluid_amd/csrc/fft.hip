// Periodic N^3 mesh transforms for the PM force convolution.
//
// The reference convolves the CIC density with tabulated vector force
// kernels through FFTW single-precision R2C/C2R transforms
// (pm.f90:112-125 per nfe^3 tile, pm.f90:288-300 / pencil_fft.f90 for the
// nc^3 coarse mesh).  Here every mesh is one periodic N^3 box in HBM and
// the transforms are three batched 1-D passes staged through LDS:
//   X : real rows  <-> N/2+1 complex (R2C/C2R via an N/2-point complex FFT)
//   Y : columns of stride N/2+1
//   Z : columns of stride N(N/2+1); the convolution multiply i*K_d and the
//       inverse Z transform are fused into the forward Z pass.
// Each workgroup transforms W adjacent columns; the 1-D FFT is a Stockham
// autosort with radix-8/4/2 butterflies held in registers and LDS used only
// for the exchange between passes.
#include "cube_internal.hpp"
#include <math.h>
#include <vector>

#define NT_FFT 256

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return cpx<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return cpx<T>{a.x - b.x, a.y - b.y}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) { return INV ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x}; }

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_small(cpx<T>* v) {
  if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cpx<T> c0 = cadd(v[0], v[2]), c1 = cadd(v[1], v[3]);
    cpx<T> d0 = csub(v[0], v[2]), d1 = mul_mi<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(c0, c1);
    v[2] = csub(c0, c1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
  } else {  // R == 8, decimation in frequency into two DFT-4
    const T h = (T)0.70710678118654752440;
    cpx<T> a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      a[k] = cadd(v[k], v[k + 4]);
      b[k] = csub(v[k], v[k + 4]);
    }
    // b_k *= w8^k, w8 = exp(-+2 pi i/8)
    b[1] = INV ? cpx<T>{h * (b[1].x - b[1].y), h * (b[1].x + b[1].y)}
               : cpx<T>{h * (b[1].x + b[1].y), h * (b[1].y - b[1].x)};
    b[2] = mul_mi<T, INV>(b[2]);
    b[3] = INV ? cpx<T>{-h * (b[3].x + b[3].y), h * (b[3].x - b[3].y)}
               : cpx<T>{h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y)};
    dft_small<T, 4, INV>(a);
    dft_small<T, 4, INV>(b);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      v[2 * m] = a[m];
      v[2 * m + 1] = b[m];
    }
  }
}

// Twiddle source: exp(-2 pi i m/L) for an L-point transform, read from a
// table of the mesh length at stride s.
template <typename T>
struct TwG {
  const cpx<T>* __restrict__ p;
  int s;
  __device__ __forceinline__ cpx<T> operator()(int m) const { return p[m * s]; }
};
// exp(-2 pi i m/16) (conjugated for the inverse) for the products m = n2*k1
// that occur inside a 16-point DFT
template <typename T, bool INV>
__device__ __forceinline__ cpx<T> w16(int m) {
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
  cpx<T> w;
  switch (m) {
    case 1: w = {c1, -s1}; break;
    case 2: w = {h, -h}; break;
    case 3: w = {s1, -c1}; break;
    case 4: w = {(T)0, (T)-1}; break;
    case 6: w = {-h, -h}; break;
    case 9: w = {-c1, s1}; break;
    default: w = {(T)1, (T)0};
  }
  if (INV) w.y = -w.y;
  return w;
}
// 16-point DFT as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
template <typename T, bool INV>
__device__ __forceinline__ void dft16(cpx<T>* v) {
  cpx<T> a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) a[n2][n1] = v[4 * n1 + n2];
    dft_small<T, 4, INV>(a[n2]);
#pragma unroll
    for (int k1 = 1; k1 < 4; k1++)
      if (n2 > 0) a[n2][k1] = cmul(a[n2][k1], w16<T, INV>(n2 * k1));
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    cpx<T> b[4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) b[n2] = a[n2][k1];
    dft_small<T, 4, INV>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = b[k2];
  }
}

// One Stockham pass over all W columns held in s[n*LD + w] (LD >= W; a
// padded LD keeps transposed loads free of LDS bank conflicts).
template <typename T, int N, int W, int LD, bool INV, int R, int NS, int NTH, typename TW>
__device__ __forceinline__ void fft_pass(cpx<T>* s, const TW& tw) {
  constexpr int B = N / R;
  constexpr int TOT = B * W;
  constexpr int PER = (TOT + NTH - 1) / NTH;
  cpx<T> v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int b = tid + q * NTH;
    if (TOT % NTH == 0 || b < TOT) {
      int w = b % W, j = b / W;
      int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) {
        cpx<T> a = s[(j + r * B) * LD + w];
        if (NS > 1 && r > 0) {
          cpx<T> t = tw(k * r * (N / (NS * R)));
          if (INV) t.y = -t.y;
          a = cmul(a, t);
        }
        v[q][r] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int b = tid + q * NTH;
    if (TOT % NTH == 0 || b < TOT) {
      if constexpr (R == 16) dft16<T, INV>(v[q]); else dft_small<T, R, INV>(v[q]);
      int w = b % W, j = b / W;
      int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; r++) s[(idxD + r * NS) * LD + w] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int W, int LD, bool INV, int NS, int NTH>
struct FftPasses {
  template <typename TW>
  __device__ __forceinline__ static void run(cpx<T>* s, const TW& tw) {
    if constexpr (NS < N) {
      constexpr int REM = N / NS;
      constexpr int R = (REM % 16 == 0) ? 16 : ((REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2));
      fft_pass<T, N, W, LD, INV, R, NS, NTH>(s, tw);
      FftPasses<T, N, W, LD, INV, NS * R, NTH>::run(s, tw);
    }
  }
};
// In-LDS FFT of W columns of length N by NTH threads; tw(m) = exp(-2 pi i m/N).
template <typename T, int N, int W, bool INV, int LD = W, int NTH = NT_FFT, typename TW>
__device__ __forceinline__ void lds_fft(cpx<T>* s, const TW& tw) {
  FftPasses<T, N, W, LD, INV, 1, NTH>::run(s, tw);
}
template <typename T, int N, int W, bool INV, int LD = W, int NTH = NT_FFT>
__device__ __forceinline__ void lds_fft(cpx<T>* s, const cpx<T>* tw, int tws) {
  FftPasses<T, N, W, LD, INV, 1, NTH>::run(s, TwG<T>{tw, tws});
}

__device__ __forceinline__ void grid_block_xy(int& bx, int& by) {
  const uint32_t L = blockIdx.x + gridDim.x * blockIdx.y, T = gridDim.x * gridDim.y;
  const uint32_t lg = xcd_remap(L, T);
  bx = (int)(lg % gridDim.x);
  by = (int)(lg / gridDim.x);
}

template <typename T>
constexpr int pick_w(int logn) {
  int e = sizeof(T) == 4 ? 8192 : 4096;
  int w = e >> logn;
  return w < 1 ? 1 : (w > 64 ? 64 : w);
}

// ------------------------------------------------------------------ X R2C
// rows of N reals (pitch N+2) -> N/2+1 complex, in place.  nrows = N*N.
template <typename T, int LOGN>
__global__ __launch_bounds__(NT_FFT) void k_fft_x_r2c(cpx<T>* data, const cpx<T>* __restrict__ tw, int64_t nrows) {
  constexpr int N = 1 << LOGN, H = N / 2, W = pick_w<T>(LOGN - 1), LD = W + 1;
  __shared__ cpx<T> s[H * LD];
  const int64_t r0 = (int64_t)blockIdx.x * W;
  {
    constexpr int PER = (H * W + NT_FFT - 1) / NT_FFT;
    cpx<T> r[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = threadIdx.x + k * NT_FFT, w = idx / H, m = idx % H;
      const int64_t row = r0 + w;
      r[k] = (idx < H * W && row < nrows) ? data[row * (H + 1) + m] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = threadIdx.x + k * NT_FFT, w = idx / H, m = idx % H;
      if (idx < H * W) s[m * LD + w] = r[k];
    }
  }
  __syncthreads();
  lds_fft<T, H, W, false, LD>(s, tw, 2);
  for (int idx = threadIdx.x; idx < (H + 1) * W; idx += NT_FFT) {
    int w = idx / (H + 1), k = idx % (H + 1);
    int64_t row = r0 + w;
    cpx<T> zk = s[(k % H) * LD + w];
    cpx<T> zc = s[((H - k) % H) * LD + w];
    zc.y = -zc.y;
    cpx<T> e{(zk.x + zc.x) * (T)0.5, (zk.y + zc.y) * (T)0.5};
    cpx<T> dd = csub(zk, zc);
    cpx<T> o{dd.y * (T)0.5, -dd.x * (T)0.5};
    cpx<T> X = cadd(e, cmul(tw[k], o));
    if (row < nrows) data[row * (H + 1) + k] = X;
  }
}

// ------------------------------------------------------------------ X C2R
// rows of N/2+1 complex (src) -> N reals at dst[row*opitch + x] (src may be
// dst: every row is read into LDS before it is written).  With f2max set,
// the other two force components F0, F1 (row pitches p0, p1) are read at the
// same nodes and max(F0^2 + F1^2 + f^2), summed in float in component
// order, goes into *f2max as float bits.
template <typename T, int LOGN>
__global__ __launch_bounds__(NT_FFT) void k_fft_x_c2r(const cpx<T>* src, T* dst, int64_t opitch,
                                                       const cpx<T>* __restrict__ tw, int64_t nrows,
                                                       const T* __restrict__ F0, int64_t p0,
                                                       const T* __restrict__ F1, int64_t p1, uint32_t* f2max) {
  constexpr int N = 1 << LOGN, H = N / 2, W = pick_w<T>(LOGN - 1), LD = W + 1;
  __shared__ cpx<T> s[H * LD];
  __shared__ cpx<T> xnyq[W];
  __shared__ float red[NT_FFT / 64];
  const int64_t r0 = (int64_t)blockIdx.x * W;
  {
    // all loads of a thread are issued before its LDS stores
    constexpr int PER = ((H + 1) * W + NT_FFT - 1) / NT_FFT;
    cpx<T> r[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NT_FFT, w = idx / (H + 1), k = idx % (H + 1);
      const int64_t row = r0 + w;
      r[q] = (idx < (H + 1) * W && row < nrows) ? src[row * (H + 1) + k] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NT_FFT, w = idx / (H + 1), k = idx % (H + 1);
      if (idx < (H + 1) * W) {
        if (k < H) s[k * LD + w] = r[q]; else xnyq[w] = r[q];
      }
    }
  }
  __syncthreads();
  // Z[k] = (X[k] + conj X[H-k]) + i exp(+2 pi i k/N) (X[k] - conj X[H-k]),
  // in place: one thread per pair (k, H-k), k = 0..H/2
  for (int idx = threadIdx.x; idx < (H / 2 + 1) * W; idx += NT_FFT) {
    int w = idx / (H / 2 + 1), k = idx % (H / 2 + 1);
    auto zof = [&](cpx<T> a, cpx<T> b, int kk) {
      b.y = -b.y;
      cpx<T> sum = cadd(a, b), dif = csub(a, b);
      cpx<T> t = tw[kk];
      t.y = -t.y;  // exp(+2 pi i kk/N)
      cpx<T> td = cmul(t, dif);
      return cpx<T>{sum.x - td.y, sum.y + td.x};  // sum + i*t*dif
    };
    cpx<T> a = s[k * LD + w];
    if (k == 0) {
      s[w] = zof(a, xnyq[w], 0);
    } else if (k == H / 2) {
      s[k * LD + w] = zof(a, a, k);
    } else {
      cpx<T> b = s[(H - k) * LD + w];
      s[k * LD + w] = zof(a, b, k);
      s[(H - k) * LD + w] = zof(b, a, H - k);
    }
  }
  __syncthreads();
  lds_fft<T, H, W, true, LD>(s, tw, 2);
  float mymax = 0.f;
  if (!f2max) {
    for (int idx = threadIdx.x; idx < H * W; idx += NT_FFT) {
      int w = idx / H, m = idx % H;
      int64_t row = r0 + w;
      if (row < nrows) *reinterpret_cast<cpx<T>*>(dst + row * opitch + 2 * m) = s[m * LD + w];
    }
  } else {
    // groups of FG nodes: their F0, F1 loads are issued together
    constexpr int PER = H * W / NT_FFT, FG = PER < 8 ? PER : 8;
    static_assert(PER % FG == 0, "whole groups");
    for (int q0 = 0; q0 < PER; q0 += FG) {
      cpx<T> a[FG], b[FG];
#pragma unroll
      for (int q = 0; q < FG; q++) {
        const int idx = threadIdx.x + (q0 + q) * NT_FFT, w = idx / H, m = idx % H;
        const int64_t row = r0 + w < nrows ? r0 + w : nrows - 1;
        a[q] = *reinterpret_cast<const cpx<T>*>(F0 + row * p0 + 2 * m);
        b[q] = *reinterpret_cast<const cpx<T>*>(F1 + row * p1 + 2 * m);
      }
#pragma unroll
      for (int q = 0; q < FG; q++) {
        const int idx = threadIdx.x + (q0 + q) * NT_FFT, w = idx / H, m = idx % H;
        const int64_t row = r0 + w;
        if (row >= nrows) continue;
        const cpx<T> z = s[m * LD + w];
        *reinterpret_cast<cpx<T>*>(dst + row * opitch + 2 * m) = z;
        float s0 = (float)a[q].x * (float)a[q].x, s1 = (float)a[q].y * (float)a[q].y;
        s0 = fmaf((float)b[q].x, (float)b[q].x, s0);
        s1 = fmaf((float)b[q].y, (float)b[q].y, s1);
        s0 = fmaf((float)z.x, (float)z.x, s0);
        s1 = fmaf((float)z.y, (float)z.y, s1);
        mymax = fmaxf(mymax, fmaxf(s0, s1));
      }
    }
  }
  if (f2max) {
    for (int off = 32; off > 0; off >>= 1) mymax = fmaxf(mymax, __shfl_xor(mymax, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mymax;
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = red[0];
      for (int i = 1; i < NT_FFT / 64; i++) m = fmaxf(m, red[i]);
      atomicMax(f2max, __float_as_uint(m));
    }
  }
}

// Load W columns of N complex values at stride `stride` into s[n*W + w]
// (columns w >= wvalid read as zero).  All loads of a thread are issued
// before any LDS store so that they are in flight together.
template <typename T, int N, int W, int NTH = NT_FFT>
__device__ __forceinline__ void col_load(cpx<T>* s, const cpx<T>* __restrict__ src, int64_t stride, int wvalid) {
  constexpr int TOT = N * W, PER = TOT / NTH;
  static_assert(TOT % NTH == 0, "whole loads per thread");
  cpx<T> r[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int idx = threadIdx.x + k * NTH, n = idx / W, w = idx % W;
    r[k] = w < wvalid ? src[n * stride + w] : cpx<T>{0, 0};
  }
#pragma unroll
  for (int k = 0; k < PER; k++) s[threadIdx.x + k * NTH] = r[k];
  __syncthreads();
}

// ----------------------------------------------------------- strided Y / Z
// AXIS 1: columns along y (stride H+1) in plane by = z.
// AXIS 2: columns along z (stride N(H+1)) at by = y; (bx, by) is the XCD-remapped block.
template <typename T, int LOGN, int AXIS, bool INV>
__global__ __launch_bounds__(NT_FFT) void k_fft_strided(cpx<T>* data, const cpx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN, KX = N / 2 + 1, W = pick_w<T>(LOGN);
  __shared__ cpx<T> s[N * W];
  int bx, by;
  grid_block_xy(bx, by);
  const int kx0 = bx * W;
  const int64_t stride = AXIS == 1 ? (int64_t)KX : (int64_t)N * KX;
  const int64_t base = AXIS == 1 ? (int64_t)by * N * KX + kx0 : (int64_t)by * KX + kx0;
  col_load<T, N, W>(s, data + base, stride, KX - kx0);
  lds_fft<T, N, W, INV>(s, tw, 1);
  for (int idx = threadIdx.x; idx < N * W; idx += NT_FFT) {
    int n = idx / W, w = idx % W;
    if (kx0 + w < KX) data[base + n * stride + w] = s[idx];
  }
}

// ------------------------------------------------------ fused Z convolution
// dst = Zinv( i*K_d . Zfwd(src) ) along z for columns (kx0.., ky = by).
// dst may be src: a workgroup reads all its columns before writing them.
// K is the imaginary part of the kernel transform over one octant,
// [d][ky'][kx][kz'] with ky'=min(ky,N-ky), kz'=min(kz,N-kz); component d is
// odd in its own axis and even in the other two.
template <typename T, int LOGN>
__global__ __launch_bounds__(NT_FFT) void k_fft_zconv(const cpx<T>* src, cpx<T>* dst, const T* __restrict__ K,
                                                       int d, const cpx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN, KX = N / 2 + 1, H = N / 2 + 1, W = pick_w<T>(LOGN);
  __shared__ cpx<T> s[N * W];
  int bx, by;
  grid_block_xy(bx, by);
  const int kx0 = bx * W;
  const int ky = by;
  const int64_t stride = (int64_t)N * KX;
  const int64_t base = (int64_t)ky * KX + kx0;
  // the kernel values of this thread's elements are loaded before the
  // forward transform, so their latency overlaps it
  constexpr int PER = N * W / NT_FFT;
  const int kyp = ky <= N / 2 ? ky : N - ky;
  const T sy = (d == 1 && ky > N / 2) ? (T)-1 : (T)1;
  T kv[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int idx = threadIdx.x + q * NT_FFT, w = idx % W, n = idx / W;
    const int kx = kx0 + w < KX ? kx0 + w : KX - 1;
    const int kzp = n <= N / 2 ? n : N - n;
    kv[q] = K[(((int64_t)d * H + kyp) * KX + kx) * H + kzp];
  }
  col_load<T, N, W>(s, src + base, stride, KX - kx0);
  lds_fft<T, N, W, false>(s, tw, 1);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int idx = threadIdx.x + q * NT_FFT, n = idx / W;
    const T sz = (d == 2 && n > N / 2) ? (T)-1 : (T)1;
    const T k = kv[q] * sy * sz;
    const cpx<T> a = s[idx];
    s[idx] = cpx<T>{-a.y * k, a.x * k};
  }
  __syncthreads();
  lds_fft<T, N, W, true>(s, tw, 1);
  for (int idx = threadIdx.x; idx < N * W; idx += NT_FFT) {
    int n = idx / W, w = idx % W;
    if (kx0 + w < KX) dst[base + n * stride + w] = s[idx];
  }
}

// ------------------------------------------------- kx-tiled complex layout
// The per-step fine-mesh convolution keeps its spectrum in a kx-tiled
// layout: element (kx, ky, kz) at ((kx/16 * N + kz) * N + ky) * 16 + kx%16,
// KT = ceil((N/2+1)/16) tiles of 16 kx (slots past N/2 are padding).  Every
// pass then moves whole cache lines: a y column strip of W kx is W*8 bytes
// of each 128-byte group (neighbouring strips complete the group), a z
// strip likewise at a stride of N*16 elements, and an x row block of W rows
// is written and read as 16*W contiguous elements per kx tile.  The R2C and
// C2R passes run out of place (real rows <-> tiled spectrum).
#define TB 16
// tiled y passes: twice the plain strip width (whole 64-byte halves of each
// group at 2048^3), with twice the threads when that doubles the LDS tile
template <typename T>
constexpr int pick_wy(int logn) { return 2 * pick_w<T>(logn) < TB ? 2 * pick_w<T>(logn) : TB; }
template <typename T>
constexpr int pick_nty(int logn) { return pick_wy<T>(logn) > pick_w<T>(logn) ? 2 * NT_FFT : NT_FFT; }
// tiled x passes: rows per workgroup, half the LDS tile of the plain x pass
template <typename T>
constexpr int pick_wtx(int logn) { return pick_w<T>(logn - 1) > 1 ? pick_w<T>(logn - 1) / 2 : 1; }
// z convolution strips: twice the y strip width (whole 64-byte halves of
// each group), with twice the threads when that doubles the LDS tile
template <typename T>
constexpr int pick_wz(int logn) { return 2 * pick_w<T>(logn) < TB ? 2 * pick_w<T>(logn) : TB; }
template <typename T>
constexpr int pick_ntz(int logn) { return pick_wz<T>(logn) > pick_w<T>(logn) ? 2 * NT_FFT : NT_FFT; }
__host__ __device__ constexpr int64_t t_tiles(int64_t N) { return (N / 2 + 1 + TB - 1) / TB; }
__host__ __device__ __forceinline__ int64_t t_index(int64_t N, int64_t kx, int64_t ky, int64_t kz) {
  return (((kx / TB) * N + kz) * N + ky) * TB + kx % TB;
}

// real rows src (pitch N+2) -> tiled spectrum dst
template <typename T, int LOGN>
__global__ __launch_bounds__(NT_FFT) void k_tx_r2c(const T* __restrict__ src, cpx<T>* __restrict__ dst,
                                                   const cpx<T>* __restrict__ tw, int64_t nrows) {
  constexpr int N = 1 << LOGN, H = N / 2, W = pick_wtx<T>(LOGN), LD = W + 1;
  constexpr int KT = (int)t_tiles(N);
  __shared__ cpx<T> s[H * LD];
  const int64_t r0 = (int64_t)blockIdx.x * W;
  {
    constexpr int PER = (H * W + NT_FFT - 1) / NT_FFT;
    cpx<T> r[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = threadIdx.x + k * NT_FFT, w = idx / H, m = idx % H;
      const int64_t row = r0 + w;
      r[k] = (idx < H * W && row < nrows) ? *reinterpret_cast<const cpx<T>*>(src + row * (N + 2) + 2 * m) : cpx<T>{0, 0};
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int idx = threadIdx.x + k * NT_FFT, w = idx / H, m = idx % H;
      if (idx < H * W) s[m * LD + w] = r[k];
    }
  }
  __syncthreads();
  lds_fft<T, H, W, false, LD>(s, tw, 2);
  // lanes run over (kx%16, row) inside one kx tile: 16*W contiguous outputs
  for (int idx = threadIdx.x; idx < KT * TB * W; idx += NT_FFT) {
    const int kl = idx % TB, w = (idx / TB) % W, kh = idx / (TB * W), k = kh * TB + kl;
    const int64_t row = r0 + w;
    if (row >= nrows) continue;
    cpx<T> X{0, 0};
    if (k <= H) {
      cpx<T> zk = s[(k % H) * LD + w];
      cpx<T> zc = s[((H - k) % H) * LD + w];
      zc.y = -zc.y;
      cpx<T> e{(zk.x + zc.x) * (T)0.5, (zk.y + zc.y) * (T)0.5};
      cpx<T> dd = csub(zk, zc);
      cpx<T> o{dd.y * (T)0.5, -dd.x * (T)0.5};
      X = cadd(e, cmul(tw[k], o));
    }
    dst[t_index(N, k, row % N, row / N)] = X;
  }
}

// tiled spectrum src -> real rows dst (pitch opitch)
template <typename T, int LOGN>
__global__ __launch_bounds__(NT_FFT) void k_tx_c2r(const cpx<T>* __restrict__ src, T* dst, int64_t opitch,
                                                   const cpx<T>* __restrict__ tw, int64_t nrows) {
  constexpr int N = 1 << LOGN, H = N / 2, W = pick_wtx<T>(LOGN), LD = W + 1;
  // s[k*LD + w] for w < W; the padding column keeps row w's Nyquist term at
  // s[w*LD + W] when W <= H, so the workgroup needs exactly H*LD entries
  constexpr int XN = W <= H ? 0 : W;
  __shared__ cpx<T> s[H * LD + XN];
  auto nyq = [](int w) { return XN ? H * LD + w : w * LD + W; };
  const int64_t r0 = (int64_t)blockIdx.x * W;
  {
    constexpr int TOT = (H / TB + 1) * TB * W, PER = (TOT + NT_FFT - 1) / NT_FFT;
    cpx<T> r[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NT_FFT, kl = idx % TB, w = (idx / TB) % W, kh = idx / (TB * W), k = kh * TB + kl;
      const int64_t row = r0 + w;
      r[q] = (idx < TOT && k <= H && row < nrows) ? src[t_index(N, k, row % N, row / N)] : cpx<T>{0, 0};
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NT_FFT, kl = idx % TB, w = (idx / TB) % W, kh = idx / (TB * W), k = kh * TB + kl;
      if (idx < TOT && k <= H) {
        if (k < H) s[k * LD + w] = r[q]; else s[nyq(w)] = r[q];
      }
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < (H / 2 + 1) * W; idx += NT_FFT) {
    int w = idx / (H / 2 + 1), k = idx % (H / 2 + 1);
    auto zof = [&](cpx<T> a, cpx<T> b, int kk) {
      b.y = -b.y;
      cpx<T> sum = cadd(a, b), dif = csub(a, b);
      cpx<T> t = tw[kk];
      t.y = -t.y;  // exp(+2 pi i kk/N)
      cpx<T> td = cmul(t, dif);
      return cpx<T>{sum.x - td.y, sum.y + td.x};  // sum + i*t*dif
    };
    cpx<T> a = s[k * LD + w];
    if (k == 0) {
      s[w] = zof(a, s[nyq(w)], 0);
    } else if (k == H / 2) {
      s[k * LD + w] = zof(a, a, k);
    } else {
      cpx<T> b = s[(H - k) * LD + w];
      s[k * LD + w] = zof(a, b, k);
      s[(H - k) * LD + w] = zof(b, a, H - k);
    }
  }
  __syncthreads();
  lds_fft<T, H, W, true, LD>(s, tw, 2);
  for (int idx = threadIdx.x; idx < H * W; idx += NT_FFT) {
    int w = idx / H, m = idx % H;
    int64_t row = r0 + w;
    if (row < nrows) *reinterpret_cast<cpx<T>*>(dst + row * opitch + 2 * m) = s[m * LD + w];
  }
}

// tiled y pass, in place: tile (kh, part, kz) = W columns kx = kh*16 +
// part*W + w along ky (stride 16)
template <typename T, int LOGN, bool INV>
__global__ __launch_bounds__(pick_nty<T>(LOGN)) void k_ty(cpx<T>* data, const cpx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN, W = pick_wy<T>(LOGN), NP = TB / W, NTH = pick_nty<T>(LOGN);
  __shared__ cpx<T> s[N * W];
  int bx, by;
  grid_block_xy(bx, by);
  const int kh = bx / NP, part = bx % NP, kz = by;
  const int64_t base = (((int64_t)kh * N + kz) * N) * TB + part * W;
  col_load<T, N, W, NTH>(s, data + base, TB, W);
  lds_fft<T, N, W, INV, W, NTH>(s, tw, 1);
  for (int idx = threadIdx.x; idx < N * W; idx += NTH) {
    const int n = idx / W, w = idx % W;
    data[base + (int64_t)n * TB + w] = s[idx];
  }
}

// tiled z convolution: dst = Zinv(i*K_d . Zfwd(src)), tile (kh, part, ky) =
// W columns along kz (stride N*16); dst may be src
template <typename T, int LOGN>
__global__ __launch_bounds__(pick_ntz<T>(LOGN)) void k_tzconv(const cpx<T>* src, cpx<T>* dst, const T* __restrict__ K, int d,
                                                   const cpx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN, KX = N / 2 + 1, H = N / 2 + 1, W = pick_wz<T>(LOGN), NP = TB / W, NTH = pick_ntz<T>(LOGN);
  constexpr int PER = N * W / NTH;
  static_assert(N * W % NTH == 0, "whole loads per thread");
  __shared__ cpx<T> s[N * W];
  int bx, by;
  grid_block_xy(bx, by);
  const int kh = bx / NP, part = bx % NP, ky = by, kx0 = kh * TB + part * W;
  const int64_t stride = (int64_t)N * TB;
  const int64_t base = ((int64_t)kh * N * N) * TB + (int64_t)ky * TB + part * W;
  const int kyp = ky <= N / 2 ? ky : N - ky;
  const T sy = (d == 1 && ky > N / 2) ? (T)-1 : (T)1;
  // K in strips of W kx: K[d][kyp][kx / W][kzp][kx % W], so the strip's
  // kernel values are one contiguous run of H*W
  constexpr int G = (int)(t_tiles(N) * TB / W);
  const T* Ks = K + (((int64_t)d * H + kyp) * G + kx0 / W) * H * W;
  col_load<T, N, W, NTH>(s, src + base, stride, W);
  lds_fft<T, N, W, false, W, NTH>(s, tw, 1);
  // the kernel values are requested only now: held across the forward FFT
  // they push the kernel past 256 VGPRs into scratch spills, which
  // cost more than this exposed load
  T kv[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int idx = threadIdx.x + q * NTH, w = idx % W, n = idx / W;
    const int kzp = n <= N / 2 ? n : N - n;
    kv[q] = Ks[kzp * W + w];
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int idx = threadIdx.x + q * NTH, n = idx / W;
    const T sz = (d == 2 && n > N / 2) ? (T)-1 : (T)1;
    const T k = kv[q] * sy * sz;
    const cpx<T> a = s[idx];
    s[idx] = cpx<T>{-a.y * k, a.x * k};
  }
  __syncthreads();
  lds_fft<T, N, W, true, W, NTH>(s, tw, 1);
  for (int idx = threadIdx.x; idx < N * W; idx += NTH) {
    const int n = idx / W, w = idx % W;
    dst[base + (int64_t)n * stride + w] = s[idx];
  }
}

// all three components from one read of the spectrum: the strip's forward z
// transform is kept in LDS while each component is multiplied by i*K_d,
// transformed back and stored to dst[d] (dst[2] may be src).  Strips are
// half as wide as k_tzconv's, so both copies fit; every column's
// arithmetic is k_tzconv's, so the outputs are the same bits.
template <typename T>
constexpr int pick_wz3(int logn) { return pick_wz<T>(logn) / 2 > 0 ? pick_wz<T>(logn) / 2 : 1; }
// meshes whose strips give every thread whole elements (all but 16^3)
template <typename T>
constexpr bool fits_z3(int logn) {
  return ((1 << logn) * pick_wz3<T>(logn)) % pick_ntz<T>(logn) == 0 && (1 << logn) * pick_wz3<T>(logn) >= pick_ntz<T>(logn);
}
template <typename T, int LOGN>
__global__ __launch_bounds__(pick_ntz<T>(LOGN)) void k_tzconv3(const cpx<T>* src, cpx<T>* dst0, cpx<T>* dst1, cpx<T>* dst2,
                                                     const T* __restrict__ K, const cpx<T>* __restrict__ tw) {
  constexpr int N = 1 << LOGN, H = N / 2 + 1, W = pick_wz3<T>(LOGN), NP = TB / W, NTH = pick_ntz<T>(LOGN);
  constexpr int WK = pick_wz<T>(LOGN);  // the kernel table's strip width
  constexpr int PER = N * W / NTH;
  static_assert(N * W % NTH == 0, "whole loads per thread");
  static_assert(WK % W == 0, "strips inside the kernel table's strips");
  __shared__ cpx<T> sa[N * W];
  __shared__ cpx<T> sb[N * W];
  int bx, by;
  grid_block_xy(bx, by);
  const int kh = bx / NP, part = bx % NP, ky = by, kx0 = kh * TB + part * W;
  const int64_t stride = (int64_t)N * TB;
  const int64_t base = ((int64_t)kh * N * N) * TB + (int64_t)ky * TB + part * W;
  const int kyp = ky <= N / 2 ? ky : N - ky;
  constexpr int G = (int)(t_tiles(N) * TB / WK);
  col_load<T, N, W, NTH>(sa, src + base, stride, W);
  lds_fft<T, N, W, false, W, NTH>(sa, tw, 1);
  cpx<T>* const dsts[3] = {dst0, dst1, dst2};
#pragma unroll 1
  for (int d = 0; d < 3; d++) {
    const T sy = (d == 1 && ky > N / 2) ? (T)-1 : (T)1;
    // K[d][kyp][kx / WK][kzp][kx % WK]
    const T* Ks = K + (((int64_t)d * H + kyp) * G + kx0 / WK) * H * WK + kx0 % WK;
    T kv[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NTH, w = idx % W, n = idx / W;
      const int kzp = n <= N / 2 ? n : N - n;
      kv[q] = Ks[kzp * WK + w];
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int idx = threadIdx.x + q * NTH, n = idx / W;
      const T sz = (d == 2 && n > N / 2) ? (T)-1 : (T)1;
      const T k = kv[q] * sy * sz;
      const cpx<T> a = sa[idx];
      sb[idx] = cpx<T>{-a.y * k, a.x * k};
    }
    __syncthreads();
    lds_fft<T, N, W, true, W, NTH>(sb, tw, 1);
    cpx<T>* const dst = dsts[d];
    for (int idx = threadIdx.x; idx < N * W; idx += NTH) {
      const int n = idx / W, w = idx % W;
      dst[base + (int64_t)n * stride + w] = sb[idx];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ dispatchers
#define LOGN_SWITCH(logn, ...)             \
  switch (logn) {                           \
    case 4: { constexpr int L = 4; __VA_ARGS__; } break;   \
    case 5: { constexpr int L = 5; __VA_ARGS__; } break;   \
    case 6: { constexpr int L = 6; __VA_ARGS__; } break;   \
    case 7: { constexpr int L = 7; __VA_ARGS__; } break;   \
    case 8: { constexpr int L = 8; __VA_ARGS__; } break;   \
    case 9: { constexpr int L = 9; __VA_ARGS__; } break;   \
    case 10: { constexpr int L = 10; __VA_ARGS__; } break; \
    case 11: { constexpr int L = 11; __VA_ARGS__; } break; \
    default: return cube_fail(h, CUBE_ERR_ARG, "mesh size 2^%d unsupported", logn); \
  }

static int ilog2(int64_t n) { int l = 0; while ((1LL << l) < n) l++; return l; }

template <typename T>
static int fwd_xy_t(cube_handle* h, Mesh& m) {
  int logn = ilog2(m.N);
  int64_t N = m.N, nrows = N * N;
  cpx<T>* c = (cpx<T>*)m.R;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  const bool fine = &m == &h->fine;
  LOGN_SWITCH(logn, {
    constexpr int WX = pick_w<T>(L - 1);
    int p1 = prof_begin(h, fine ? PS_FFT_XF : PS_FFT_C);
    k_fft_x_r2c<T, L><<<dim3((unsigned)((nrows + WX - 1) / WX)), NT_FFT, 0, h->stream>>>(c, tw, nrows);
    prof_end(h, p1);
    constexpr int WS = pick_w<T>(L);
    int p2 = prof_begin(h, fine ? PS_FFT_YF : PS_FFT_C);
    k_fft_strided<T, L, 1, false><<<dim3((unsigned)((N / 2 + 1 + WS - 1) / WS), (unsigned)N), NT_FFT, 0, h->stream>>>(c, tw);
    prof_end(h, p2);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_fwd_xy(cube_handle* h, Mesh& m) {
  return m.prec == 4 ? fwd_xy_t<float>(h, m) : fwd_xy_t<double>(h, m);
}

template <typename T>
static int fft3_r2c_t(cube_handle* h, Mesh& m) {
  int r = fwd_xy_t<T>(h, m);
  if (r) return r;
  int logn = ilog2(m.N);
  int64_t N = m.N;
  cpx<T>* c = (cpx<T>*)m.R;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  LOGN_SWITCH(logn, {
    constexpr int WS = pick_w<T>(L);
    k_fft_strided<T, L, 2, false><<<dim3((unsigned)((N / 2 + 1 + WS - 1) / WS), (unsigned)N), NT_FFT, 0, h->stream>>>(c, tw);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_fft3_r2c(cube_handle* h, Mesh& m) {
  return m.prec == 4 ? fft3_r2c_t<float>(h, m) : fft3_r2c_t<double>(h, m);
}

template <typename T>
static int zconv_t(cube_handle* h, Mesh& m, int d, void* dst) {
  int logn = ilog2(m.N);
  int64_t N = m.N;
  int p1 = prof_begin(h, &m == &h->fine ? PS_FFT_ZCONV : PS_FFT_C);
  LOGN_SWITCH(logn, {
    constexpr int WS = pick_w<T>(L);
    k_fft_zconv<T, L><<<dim3((unsigned)((N / 2 + 1 + WS - 1) / WS), (unsigned)N), NT_FFT, 0, h->stream>>>(
        (const cpx<T>*)m.R, (cpx<T>*)dst, (const T*)m.K, d, (const cpx<T>*)m.tw);
  });
  prof_end(h, p1);
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_zconv(cube_handle* h, Mesh& m, int d, void* dst) {
  if (dst != m.R && dst != m.W) return cube_fail(h, CUBE_ERR_ARG, "z convolution target must be R or W");
  return m.prec == 4 ? zconv_t<float>(h, m, d, dst) : zconv_t<double>(h, m, d, dst);
}

template <typename T>
static int inv_yx_t(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch, const void* F0, int64_t p0,
                    const void* F1, int64_t p1, uint32_t* f2max) {
  int logn = ilog2(m.N);
  int64_t N = m.N, nrows = N * N;
  cpx<T>* c = (cpx<T>*)buf;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  const bool fine = &m == &h->fine;
  LOGN_SWITCH(logn, {
    constexpr int WS = pick_w<T>(L);
    int pa = prof_begin(h, fine ? PS_FFT_YI : PS_FFT_C);
    k_fft_strided<T, L, 1, true><<<dim3((unsigned)((N / 2 + 1 + WS - 1) / WS), (unsigned)N), NT_FFT, 0, h->stream>>>(c, tw);
    prof_end(h, pa);
    constexpr int WX = pick_w<T>(L - 1);
    int pb = prof_begin(h, fine ? PS_FFT_XI : PS_FFT_C);
    k_fft_x_c2r<T, L><<<dim3((unsigned)((nrows + WX - 1) / WX)), NT_FFT, 0, h->stream>>>(
        c, (T*)out, opitch, tw, nrows, (const T*)F0, p0, (const T*)F1, p1, f2max);
    prof_end(h, pb);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_inv_yx(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch, const void* F0, int64_t p0,
                const void* F1, int64_t p1, uint32_t* f2max) {
  if (buf != m.R && buf != m.W) return cube_fail(h, CUBE_ERR_ARG, "inverse transform source must be R or W");
  if (out == buf && opitch != m.N + 2) return cube_fail(h, CUBE_ERR_ARG, "in-place C2R needs pitch N+2");
  if (f2max && (!F0 || !F1)) return cube_fail(h, CUBE_ERR_ARG, "force maximum needs both other components");
  return m.prec == 4 ? inv_yx_t<float>(h, m, buf, out, opitch, F0, p0, F1, p1, f2max)
                     : inv_yx_t<double>(h, m, buf, out, opitch, F0, p0, F1, p1, f2max);
}

// ---------------------------------------- tiled fine-mesh convolution host
template <typename T>
static int t_fwd_t(cube_handle* h, Mesh& m, const void* src, void* dst) {
  int logn = ilog2(m.N);
  int64_t N = m.N, nrows = N * N;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  LOGN_SWITCH(logn, {
    constexpr int WX = pick_wtx<T>(L);
    int p1 = prof_begin(h, PS_FFT_XF);
    k_tx_r2c<T, L><<<dim3((unsigned)((nrows + WX - 1) / WX)), NT_FFT, 0, h->stream>>>((const T*)src, (cpx<T>*)dst, tw, nrows);
    prof_end(h, p1);
    constexpr int WY = pick_wy<T>(L);
    int p2 = prof_begin(h, PS_FFT_YF);
    k_ty<T, L, false><<<dim3((unsigned)(t_tiles(N) * (TB / WY)), (unsigned)N), pick_nty<T>(L), 0, h->stream>>>((cpx<T>*)dst, tw);
    prof_end(h, p2);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_t_fwd(cube_handle* h, Mesh& m, const void* src, void* dst) {
  return m.prec == 4 ? t_fwd_t<float>(h, m, src, dst) : t_fwd_t<double>(h, m, src, dst);
}

template <typename T>
static int t_zconv_t(cube_handle* h, Mesh& m, int d, const void* src, void* dst) {
  int logn = ilog2(m.N);
  int64_t N = m.N;
  int p1 = prof_begin(h, PS_FFT_ZCONV);
  LOGN_SWITCH(logn, {
    constexpr int WZ = pick_wz<T>(L);
    k_tzconv<T, L><<<dim3((unsigned)(t_tiles(N) * (TB / WZ)), (unsigned)N), pick_ntz<T>(L), 0, h->stream>>>(
        (const cpx<T>*)src, (cpx<T>*)dst, (const T*)m.K, d, (const cpx<T>*)m.tw);
  });
  prof_end(h, p1);
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_t_zconv(cube_handle* h, Mesh& m, int d, const void* src, void* dst) {
  return m.prec == 4 ? t_zconv_t<float>(h, m, d, src, dst) : t_zconv_t<double>(h, m, d, src, dst);
}
bool mesh_t_zconv3_fits(const Mesh& m) {
  if (m.prec != 4) return false;
  const int logn = ilog2(m.N);
  bool ok = false;
  switch (logn) {
    case 4: ok = fits_z3<float>(4); break;
    case 5: ok = fits_z3<float>(5); break;
    case 6: ok = fits_z3<float>(6); break;
    case 7: ok = fits_z3<float>(7); break;
    case 8: ok = fits_z3<float>(8); break;
    case 9: ok = fits_z3<float>(9); break;
    case 10: ok = fits_z3<float>(10); break;
    case 11: ok = fits_z3<float>(11); break;
    default: ok = false;
  }
  return ok;
}
int mesh_t_zconv3(cube_handle* h, Mesh& m, const void* src, void* const dst[3]) {
  if (m.prec != 4) return cube_fail(h, CUBE_ERR_ARG, "three-component z convolution is built for FP32 meshes");
  int logn = ilog2(m.N);
  int64_t N = m.N;
  int p1 = prof_begin(h, PS_FFT_ZCONV);
  LOGN_SWITCH(logn, {
    if constexpr (fits_z3<float>(L)) {
      constexpr int WZ = pick_wz3<float>(L);
      k_tzconv3<float, L><<<dim3((unsigned)(t_tiles(N) * (TB / WZ)), (unsigned)N), pick_ntz<float>(L), 0, h->stream>>>(
          (const cpx<float>*)src, (cpx<float>*)dst[0], (cpx<float>*)dst[1], (cpx<float>*)dst[2], (const float*)m.K,
          (const cpx<float>*)m.tw);
    } else {
      return cube_fail(h, CUBE_ERR_ARG, "mesh 2^%d too small for the three-component z convolution", L);
    }
  });
  prof_end(h, p1);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

template <typename T>
static int t_inv_t(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch) {
  int logn = ilog2(m.N);
  int64_t N = m.N, nrows = N * N;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  LOGN_SWITCH(logn, {
    constexpr int WY = pick_wy<T>(L);
    int pa = prof_begin(h, PS_FFT_YI);
    k_ty<T, L, true><<<dim3((unsigned)(t_tiles(N) * (TB / WY)), (unsigned)N), pick_nty<T>(L), 0, h->stream>>>((cpx<T>*)buf, tw);
    prof_end(h, pa);
    constexpr int WX = pick_wtx<T>(L);
    int pb = prof_begin(h, PS_FFT_XI);
    k_tx_c2r<T, L><<<dim3((unsigned)((nrows + WX - 1) / WX)), NT_FFT, 0, h->stream>>>((const cpx<T>*)buf, (T*)out, opitch, tw,
                                                                                     nrows);
    prof_end(h, pb);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_t_inv(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch) {
  if (out == buf) return cube_fail(h, CUBE_ERR_ARG, "tiled C2R runs out of place");
  return m.prec == 4 ? t_inv_t<float>(h, m, buf, out, opitch) : t_inv_t<double>(h, m, buf, out, opitch);
}

// -------------------------------------------------------------- allocation
int mesh_alloc(cube_handle* h, Mesh& m, int64_t N, int prec, bool with_S, bool with_K) {
  m.N = N;
  m.prec = prec;
  int64_t H = N / 2 + 1;
  HIP_TRY(h, hipMalloc(&m.R, m.bytes_big()));
  HIP_TRY(h, hipMalloc(&m.W, m.bytes_big()));
  if (with_S) {
    HIP_TRY(h, hipMalloc(&m.S, (size_t)prec * N * N * N));
    HIP_TRY(h, hipMalloc(&m.X, m.bytes_big()));
  }
  // kx padded to whole 16-wide tiles: the fine mesh's strip layout needs it
  if (with_K) HIP_TRY(h, hipMalloc(&m.K, (size_t)3 * H * t_tiles(N) * TB * H * prec));
  HIP_TRY(h, hipMalloc(&m.tw, (size_t)N * 2 * prec));
  std::vector<double> tw(2 * N);
  for (int64_t j = 0; j < N; j++) {
    double a = -2.0 * M_PI * (double)j / (double)N;
    tw[2 * j] = cos(a);
    tw[2 * j + 1] = sin(a);
  }
  if (prec == 8) {
    HIP_TRY(h, hipMemcpy(m.tw, tw.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice));
  } else {
    std::vector<float> twf(2 * N);
    for (int64_t j = 0; j < 2 * N; j++) twf[j] = (float)tw[j];
    HIP_TRY(h, hipMemcpy(m.tw, twf.data(), sizeof(float) * 2 * N, hipMemcpyHostToDevice));
  }
  return 0;
}
void mesh_free(Mesh& m) {
  if (m.R) (void)hipFree(m.R);
  if (m.W) (void)hipFree(m.W);
  if (m.S) (void)hipFree(m.S);
  if (m.X) (void)hipFree(m.X);
  if (m.K) (void)hipFree(m.K);
  if (m.tw) (void)hipFree(m.tw);
  m = Mesh();
}

// ---------------------------------------------------------- fine kernel
// Im of the DFT of the tabulated fine kernel (kernel_f.f90:14-37) evaluated
// directly on the N^3 mesh as three separable 16-term contractions in FP64:
//   Im K_d(k) = -sum_a prod_e g_de(k_e, a_e) K_d(a),  a in [0,16)^3,
//   g_de = 2 sin(2 pi k a/N) for e == d, w(a) cos(2 pi k a/N) otherwise
//   (w(0)=1, w(a>0)=2), which is the odd/even mirror of the table.
__device__ __forceinline__ double g_fac(const double* cs, const double* sn, int64_t N, int e, int d, int k, int a) {
  int64_t m = ((int64_t)k * a) % N;
  if (e == d) return 2.0 * sn[m];
  return (a == 0 ? 1.0 : 2.0) * cs[m];
}
__global__ void k_kf_stage1(const float* __restrict__ tab, int pp_on, int pp_range, const double* cs,
                            const double* sn, int64_t N, int64_t H, double* A) {
  // A[d][kx][y][z] = sum_x K_d(x,y,z) g_d0(kx,x)
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t tot = 3 * H * 256;
  if (i >= tot) return;
  int z = (int)(i % 16), y = (int)((i / 16) % 16);
  int64_t kx = (i / 256) % H;
  int d = (int)(i / (256 * H));
  double s = 0;
  for (int x = 0; x < 16; x++) {
    double kv = tab[((z * 16 + y) * 16 + x) * 3 + d];
    if (pp_on && x <= pp_range && y <= pp_range && z <= pp_range) kv = 0;
    s += kv * g_fac(cs, sn, N, 0, d, (int)kx, x);
  }
  A[i] = s;
}
__global__ void k_kf_stage2(const double* A, const double* cs, const double* sn, int64_t N, int64_t H, double* B) {
  // B[d][kx][ky][z] = sum_y A[d][kx][y][z] g_d1(ky,y)
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t tot = 3 * H * H * 16;
  if (i >= tot) return;
  int z = (int)(i % 16);
  int64_t ky = (i / 16) % H, kx = (i / (16 * H)) % H;
  int d = (int)(i / (16 * H * H));
  double s = 0;
  for (int y = 0; y < 16; y++) s += A[((d * H + kx) * 16 + y) * 16 + z] * g_fac(cs, sn, N, 1, d, (int)ky, y);
  B[i] = s;
}
template <typename T>
__global__ void k_kf_stage3(const double* B, const double* cs, const double* sn, int64_t N, int64_t H, int W, double scale,
                            T* K) {
  // K[d][ky][kx / W][kz][kx % W] = -scale * sum_z B[d][kx][ky][z] g_d2(kz,z)
  // over kx padded to whole 16-wide tiles (the padding repeats kx = H-1)
  const int64_t G = t_tiles(N) * TB / W;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t tot = 3 * H * G * W * H;
  if (i >= tot) return;
  const int64_t w = i % W, kz = (i / W) % H, g = (i / (W * H)) % G, ky = (i / (W * H * G)) % H;
  const int64_t kx = g * W + w < H ? g * W + w : H - 1;
  int d = (int)(i / (W * H * G * H));
  const double* b = B + ((d * H + kx) * H + ky) * 16;
  double s = 0;
  for (int z = 0; z < 16; z++) s += b[z] * g_fac(cs, sn, N, 2, d, (int)kz, z);
  K[i] = (T)(-scale * s);
}

int fine_kernel_build(cube_handle* h, const float* table) {
  Mesh& m = h->fine;
  int64_t N = m.N, H = N / 2 + 1;
  std::vector<double> cs(N), sn(N);
  for (int64_t j = 0; j < N; j++) {
    cs[j] = cos(2.0 * M_PI * (double)j / (double)N);
    sn[j] = sin(2.0 * M_PI * (double)j / (double)N);
  }
  float* dtab; double *dcs, *dsn, *A, *B;
  HIP_TRY(h, hipMalloc(&dtab, sizeof(float) * 4096 * 3));
  HIP_TRY(h, hipMalloc(&dcs, sizeof(double) * N));
  HIP_TRY(h, hipMalloc(&dsn, sizeof(double) * N));
  HIP_TRY(h, hipMalloc(&A, sizeof(double) * 3 * H * 256));
  HIP_TRY(h, hipMalloc(&B, sizeof(double) * 3 * H * H * 16));
  HIP_TRY(h, hipMemcpy(dtab, table, sizeof(float) * 4096 * 3, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(dcs, cs.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(dsn, sn.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  // the strip width of the tiled z convolution (k_tzconv) for this mesh
  const int W = m.prec == 4 ? pick_wz<float>(ilog2(N)) : pick_wz<double>(ilog2(N));
  int64_t t1 = 3 * H * 256, t2 = 3 * H * H * 16, t3 = 3 * H * t_tiles(N) * TB * H;
  k_kf_stage1<<<(unsigned)((t1 + 255) / 256), 256, 0, h->stream>>>(dtab, h->p.pp_on, h->p.pp_range, dcs, dsn, N, H, A);
  k_kf_stage2<<<(unsigned)((t2 + 255) / 256), 256, 0, h->stream>>>(A, dcs, dsn, N, H, B);
  double scale = 1.0 / ((double)N * N * N);
  if (m.prec == 4)
    k_kf_stage3<float><<<(unsigned)((t3 + 255) / 256), 256, 0, h->stream>>>(B, dcs, dsn, N, H, W, scale, (float*)m.K);
  else
    k_kf_stage3<double><<<(unsigned)((t3 + 255) / 256), 256, 0, h->stream>>>(B, dcs, dsn, N, H, W, scale, (double*)m.K);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  (void)hipFree(dtab); (void)hipFree(dcs); (void)hipFree(dsn); (void)hipFree(A); (void)hipFree(B);
  return 0;
}

// ---------------------------------------------------------- coarse kernel
// kernel_c.f90:16-117 with one image: ck = -r/r^3 (r in fine-cell units),
// the 4^3 table of wfxyzc.2.ascii copied into the eight corner octants with
// odd/even signs, R2C in FP64, then LRCKCORR for |k|<=8, k_d != 0.
__global__ void k_ck_fill(double* R, int64_t N, int d, int corrected, const float* __restrict__ tab) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N * N * N) return;
  int64_t x = i % N, y = (i / N) % N, z = i / (N * N);
  int64_t g[3] = {x, y, z};
  double rx[3];
  for (int e = 0; e < 3; e++) rx[e] = (double)(CUBE_NCELL * (((g[e] + N / 2) % N) - N / 2));
  double r = sqrt(rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]);
  double v = r == 0 ? 0.0 : -rx[d] / (r * r * r);
  if (corrected) {
    int t[3];
    bool mir[3], in = true;
    for (int e = 0; e < 3; e++) {
      if (g[e] <= 3) { t[e] = (int)g[e]; mir[e] = false; }
      else if (g[e] >= N - 3) { t[e] = (int)(N - g[e]); mir[e] = true; }
      else in = false;
    }
    if (in) v = (mir[d] ? -1.0 : 1.0) * (double)tab[((t[2] * 16 + t[1] * 4 + t[0]) * 3) + d];
  }
  R[(z * N + y) * (N + 2) + x] = v;
}
// octant extraction of Im of the R2C output
__global__ void k_ck_extract(const double* C, int64_t N, double* out) {
  int64_t H = N / 2 + 1;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= H * H * H) return;
  int64_t kz = i % H, kx = (i / H) % H, ky = i / (H * H);
  out[i] = C[((kz * N + ky) * H + kx) * 2 + 1];
}
template <typename T>
__global__ void k_ck_correct(const double* kc, const double* ku, int64_t N, int d, T* K) {
  int64_t H = N / 2 + 1;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= H * H * H) return;
  int64_t kz = i % H, kx = (i / H) % H, ky = i / (H * H);
  int64_t kk[3] = {kx, ky, kz};
  double kv[3], ks[3];
  for (int e = 0; e < 3; e++) {
    kv[e] = (double)(((kk[e] + N / 2) % N) - N / 2);  // mod(i+nc/2-1,nc)-nc/2 with 1-based i
    ks[e] = 2.0 * sin(M_PI * kv[e] / (double)N);
  }
  double kr = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
  double v = kc[i];
  if (!(kr > 8.0 || kv[d] == 0)) {
    if (kk[d] == N / 2) v = 0.0;  // odd component at its own Nyquist plane: exactly zero
    else v = v * 0.25 * M_PI * ks[d] / (ks[0] * ks[0] + ks[1] * ks[1] + ks[2] * ks[2]) / ku[i];
  }
  K[(int64_t)d * H * H * H + i] = (T)(v / ((double)N * N * N));
}

int coarse_kernel_build(cube_handle* h, const float* table) {
  Mesh tmp;
  int64_t N = h->coarse.N, H = N / 2 + 1, H3 = H * H * H;
  int r = mesh_alloc(h, tmp, N, 8, false, false);
  if (r) return r;
  float* dtab;
  double *kc, *ku;
  HIP_TRY(h, hipMalloc(&dtab, sizeof(float) * 64 * 3));
  HIP_TRY(h, hipMalloc(&kc, sizeof(double) * H3));
  HIP_TRY(h, hipMalloc(&ku, sizeof(double) * H3));
  HIP_TRY(h, hipMemcpy(dtab, table, sizeof(float) * 64 * 3, hipMemcpyHostToDevice));
  unsigned gfill = (unsigned)((N * N * N + 255) / 256), gext = (unsigned)((H3 + 255) / 256);
  for (int d = 0; d < 3; d++) {
    k_ck_fill<<<gfill, 256, 0, h->stream>>>((double*)tmp.R, N, d, 1, dtab);
    if ((r = mesh_fft3_r2c(h, tmp))) return r;
    k_ck_extract<<<gext, 256, 0, h->stream>>>((const double*)tmp.R, N, kc);
    k_ck_fill<<<gfill, 256, 0, h->stream>>>((double*)tmp.R, N, d, 0, dtab);
    if ((r = mesh_fft3_r2c(h, tmp))) return r;
    k_ck_extract<<<gext, 256, 0, h->stream>>>((const double*)tmp.R, N, ku);
    if (h->coarse.prec == 4)
      k_ck_correct<float><<<gext, 256, 0, h->stream>>>(kc, ku, N, d, (float*)h->coarse.K);
    else
      k_ck_correct<double><<<gext, 256, 0, h->stream>>>(kc, ku, N, d, (double*)h->coarse.K);
    HIP_TRY(h, hipGetLastError());
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  (void)hipFree(dtab); (void)hipFree(kc); (void)hipFree(ku);
  mesh_free(tmp);
  return 0;
}

// ------------------------------------------------------ generic passes
// One strided pass (axis 1: y, axis 2: z) over a complex buffer of mesh m.
template <typename T>
static int pass_t(cube_handle* h, Mesh& m, void* buf, int axis, bool inv) {
  int logn = ilog2(m.N);
  int64_t N = m.N;
  cpx<T>* c = (cpx<T>*)buf;
  const cpx<T>* tw = (const cpx<T>*)m.tw;
  LOGN_SWITCH(logn, {
    constexpr int WS = pick_w<T>(L);
    dim3 g((unsigned)((N / 2 + 1 + WS - 1) / WS), (unsigned)N);
    if (axis == 1 && !inv) k_fft_strided<T, L, 1, false><<<g, NT_FFT, 0, h->stream>>>(c, tw);
    if (axis == 1 && inv) k_fft_strided<T, L, 1, true><<<g, NT_FFT, 0, h->stream>>>(c, tw);
    if (axis == 2 && !inv) k_fft_strided<T, L, 2, false><<<g, NT_FFT, 0, h->stream>>>(c, tw);
    if (axis == 2 && inv) k_fft_strided<T, L, 2, true><<<g, NT_FFT, 0, h->stream>>>(c, tw);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_pass(cube_handle* h, Mesh& m, void* buf, int axis, bool inverse) {
  return m.prec == 4 ? pass_t<float>(h, m, buf, axis, inverse) : pass_t<double>(h, m, buf, axis, inverse);
}
// C2R along x of a complex buffer into real rows of pitch opitch (no force maximum).
template <typename T>
static int c2r_t(cube_handle* h, Mesh& m, const void* src, void* dst, int64_t opitch) {
  int logn = ilog2(m.N);
  int64_t N = m.N, nrows = N * N;
  LOGN_SWITCH(logn, {
    constexpr int WX = pick_w<T>(L - 1);
    k_fft_x_c2r<T, L><<<dim3((unsigned)((nrows + WX - 1) / WX)), NT_FFT, 0, h->stream>>>(
        (const cpx<T>*)src, (T*)dst, opitch, (const cpx<T>*)m.tw, nrows, nullptr, 0, nullptr, 0, nullptr);
  });
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int mesh_c2r(cube_handle* h, Mesh& m, const void* src, void* dst, int64_t opitch) {
  return m.prec == 4 ? c2r_t<float>(h, m, src, dst, opitch) : c2r_t<double>(h, m, src, dst, opitch);
}
