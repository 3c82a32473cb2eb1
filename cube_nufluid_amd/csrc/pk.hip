// Matter power spectrum of the resident particles on the fine mesh
// (utilities/cicpower.f90:85-257 density, utilities/powerspectrum.f90:27-137
// estimator).
//
// Density: CIC of every particle on the nf^3 fine mesh.  cicpower puts the
// nodes at cell centres (pos-0.5) while the PM density has them at cell left
// edges; the two differ by a uniform half-cell shift, which changes only
// the phase of each Fourier mode, so |delta_k|^2 is the same.  delta =
// rho/mean - 1 and the forward transform is divided by nf^3
// (powerspectrum.f90:40-46).  Modes: the half space kx >= 0 with the
// conjugate duplicates on the kx = 0 and kx = nf/2 planes skipped as in
// powerspectrum.f90:63-66; signed frequencies mod(i+nyq-1,ng)-nyq; linear
// bins nint(|k|) with |k| in real(4) as the reference computes it;
// Delta^2 = |delta_k|^2 / sinc^4 * 4 pi |k|^3 with the CIC window
// sinc = prod_d sin(pi k_d/ng)/(pi k_d/ng).
#include "cube_internal.hpp"
#include <math.h>
#include <mutex>

#define PK_NT 256
#define PK_MAXBIN 2048

template <typename T>
__global__ __launch_bounds__(PK_NT) void k_pk_bin(const cpx<T>* __restrict__ c, int64_t N, double inv_norm, int nbin,
                                                  double* bins) {
  __shared__ double sb[3 * (PK_MAXBIN + 1)];
  for (int i = threadIdx.x; i < 3 * (nbin + 1); i += PK_NT) sb[i] = 0;
  __syncthreads();
  const int64_t H = N / 2 + 1, tot = H * N * N, nyq = N / 2;
  for (int64_t i = (int64_t)blockIdx.x * PK_NT + threadIdx.x; i < tot; i += (int64_t)gridDim.x * PK_NT) {
    const int64_t kx = i % H, ky = (i / H) % N, kz = i / (H * N);
    if (kx == 0 && ky == 0 && kz == 0) continue;
    const bool edge = kx == 0 || kx == nyq;
    if (edge && ky > nyq) continue;
    if (edge && (ky == 0 || ky == nyq) && kz > nyq) continue;
    const int64_t q[3] = {(kx + nyq) % N - nyq, (ky + nyq) % N - nyq, (kz + nyq) % N - nyq};
    const float kr = sqrtf((float)(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
    const int ib = (int)floorf(kr + 0.5f);
    if (ib > nbin) continue;
    double sinc = 1;
    for (int d = 0; d < 3; d++)
      if (q[d]) {
        const double x = M_PI * (double)q[d] / (double)N;
        sinc *= sin(x) / x;
      }
    const cpx<T> v = c[i];
    const double re = (double)v.x * inv_norm, im = (double)v.y * inv_norm;
    const double s2 = sinc * sinc;
    const double amp = (re * re + im * im) / (s2 * s2) * 4.0 * M_PI * (double)kr * (double)kr * (double)kr;
    atomicAdd(&sb[3 * ib], 1.0);
    atomicAdd(&sb[3 * ib + 1], (double)kr);
    atomicAdd(&sb[3 * ib + 2], amp);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * (nbin + 1); i += PK_NT)
    if (sb[i] != 0) atomicAdd(&bins[i], sb[i]);
}

// Scratch of cube_power_spectrum_noalloc: the per-bin partial sums (mode
// count, sum of |k|, sum of Delta^2 for bins 0..PK_MAXBIN), 48 KiB.  Every
// GPU the library runs on holds its own copy until the process ends.  In
// host code `scratch` only names the symbol; its device address on each
// GPU comes from hipGetSymbolAddress at cube_init (pk_scratch_setup).
static __device__ __align__(256) double scratch[3 * (PK_MAXBIN + 1)];
#define PK_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[PK_MAX_DEVICES];

int pk_scratch_setup(cube_handle* h) {
  const int dev = h->p.device;
  if (dev < 0 || dev >= PK_MAX_DEVICES) return cube_fail(h, CUBE_ERR_ARG, "device %d outside [0, %d)", dev, PK_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return cube_fail(h, CUBE_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
  g_scratch_ptr[dev] = (double*)ptr;
  return 0;
}

__global__ void k_pk_zero(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Shared body of both entry points: argument checks, CIC density, R2C,
// binning into `bins` (3 doubles per bin 0..nbin, cleared here by a fill
// kernel), and the strided copies of bins 1..nbin into the caller's arrays.
static int pk_run(cube_handle* h, int nbin, double* modes, double* ksum, double* d2sum, double* bins) {
  const int64_t N = h->fine.N;
  const int want = (int)lrint((double)(N / 2) * sqrt(3.0));
  if (nbin != want)
    return cube_fail(h, CUBE_ERR_ARG, "nbin=%d; the linear binning of a %lld^3 mesh has nint(nyquist*sqrt(3))=%d bins", nbin,
                     (long long)N, want);
  if (nbin > PK_MAXBIN) return cube_fail(h, CUBE_ERR_ARG, "nbin=%d above %d", nbin, PK_MAXBIN);
  double mass = 0;
  int r = cic_fine(h, &mass);
  if (r) return r;
  if ((r = mesh_fft3_r2c(h, h->fine))) return r;
  const int nv = 3 * (nbin + 1);
  k_pk_zero<<<(nv + PK_NT - 1) / PK_NT, PK_NT, 0, h->stream>>>(bins, nv);
  HIP_TRY(h, hipGetLastError());
  const double mean = mass / ((double)N * N * N);
  const double inv_norm = 1.0 / ((double)N * N * N * mean);
  const int64_t tot = (N / 2 + 1) * N * N;
  const unsigned grid = (unsigned)std::min<int64_t>((tot + PK_NT - 1) / PK_NT, 4096);
  if (h->prec == 4)
    k_pk_bin<float><<<grid, PK_NT, 0, h->stream>>>((const cpx<float>*)h->fine.R, N, inv_norm, nbin, bins);
  else
    k_pk_bin<double><<<grid, PK_NT, 0, h->stream>>>((const cpx<double>*)h->fine.R, N, inv_norm, nbin, bins);
  HIP_TRY(h, hipGetLastError());
  // the reference reports bins 1..nbin (xip=xi(:,1:)); field f of bin b is bins[3b+f]
  double* dst[3] = {modes, ksum, d2sum};
  for (int f = 0; f < 3; f++)
    HIP_TRY(h, hipMemcpy2DAsync(dst[f], sizeof(double), bins + 3 + f, 3 * sizeof(double), sizeof(double), (size_t)nbin,
                                hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return CUBE_OK;
}

extern "C" int cube_power_spectrum(cube_handle* h, int nbin, double* modes, double* ksum, double* d2sum) {
  CUBE_EXCLUSIVE;
  if (!h || !modes || !ksum || !d2sum) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  if (nbin < 1 || nbin > PK_MAXBIN) return cube_fail(h, CUBE_ERR_ARG, "nbin=%d outside [1, %d]", nbin, PK_MAXBIN);
  double* bins = nullptr;
  HIP_TRY(h, hipMalloc(&bins, sizeof(double) * 3 * (nbin + 1)));
  const int r = pk_run(h, nbin, modes, ksum, d2sum, bins);
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(bins);
  return r;
}

// The same estimator without any device allocation: the partial sums live
// in this file's static `scratch` array (at most PK_MAXBIN bins).  All calls
// on one GPU share that array, so calls that overlap on one GPU (several
// host threads, or handles on different streams) race; the bins are
// cleared at the start of every call.
extern "C" int cube_power_spectrum_noalloc(cube_handle* h, int nbin, double* modes, double* ksum, double* d2sum) {
  if (!h || !modes || !ksum || !d2sum) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  if (nbin < 1 || nbin > PK_MAXBIN)
    return cube_fail(h, CUBE_ERR_ARG, "nbin=%d outside [1, %d] (the allocation-free scratch holds %d bins)", nbin,
                     PK_MAXBIN, PK_MAXBIN);
  double* bins = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    const int dev = h->p.device;
    if (dev >= 0 && dev < PK_MAX_DEVICES) bins = g_scratch_ptr[dev];
  }
  if (!bins) return cube_fail(h, CUBE_ERR_ARG, "no scratch address for device %d (cube_init sets it)", h->p.device);
  return pk_run(h, nbin, modes, ksum, d2sum, bins);
}
