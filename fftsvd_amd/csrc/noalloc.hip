// Allocation-free panel integrals: fftsvd_integrate_noalloc.
//
// The variant of fftsvd_integrate for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  Inputs and outputs are the caller's device arrays;
// the per-panel geometry records the integral kernel reads (the scratch
// that fftsvd_integrate allocates and uploads) are built on the device
// into a file-scope array of fixed size.  Nothing here allocates, frees,
// copies or clears memory: the call is two kernel launches on the caller's
// stream.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/fftsvd_hip.h"
#include "geometry.h"

namespace fsvd {
void launch_integrate_all(int kernel, double kappa, const PanelGeom* panels, int npan, const double* pts, int npt,
                          double* slp, double* dlp, hipStream_t s);
void set_last_error(const char* msg);
}  // namespace fsvd

using namespace fsvd;

// Geometry records of at most FFTSVD_NOALLOC_MAX_PANELS panels (656 bytes
// each, 10.75 MB); every GPU the library runs on holds its own copy until
// the process ends.  Calls on one GPU share it: calls that overlap in time
// on one GPU (two streams, two host threads) race on it.
static_assert(sizeof(PanelGeom) % 8 == 0, "PanelGeom is an array of doubles");
static __device__ __align__(256) PanelGeom scratch[FFTSVD_NOALLOC_MAX_PANELS];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
PanelGeom* g_scratch[MAX_DEVICES] = {};   // device address of `scratch`, per device

// The device address of this device's copy of `scratch`, looked up once per
// device (the array's name is a host-side stand-in, not a device address).
PanelGeom* scratch_ptr(int dev) {
   if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range");
   std::lock_guard<std::mutex> lk(g_mu);
   if (!g_scratch[dev]) {
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
         throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      g_scratch[dev] = (PanelGeom*)p;
   }
   return g_scratch[dev];
}

// One panel record per thread (FlatPanel_allocate + FlatPanel_moments,
// FlatPanel.c:5-179,210-298), the same arithmetic as the host set-up.
__global__ void __launch_bounds__(64) k_panel_geometry(const double* __restrict__ verts9, int npan,
                                                       PanelGeom* __restrict__ out) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= npan) return;
   const double* v = verts9 + 9 * (long long)i;
   double v0[3], v1[3], v2[3];
   for (int d = 0; d < 3; d++) {
      v0[d] = v[d];
      v1[d] = v[3 + d];
      v2[d] = v[6 + d];
   }
   PanelGeom g;
   geom::panel_geometry_hd(v0, v1, v2, &g);
   const double* src = reinterpret_cast<const double*>(&g);
   double* dst = reinterpret_cast<double*>(out + i);
   for (int k = 0; k < (int)(sizeof(PanelGeom) / sizeof(double)); k++) dst[k] = src[k];
}

}  // namespace

extern "C" int fftsvd_integrate_noalloc(int kernel, double kappa, const double* verts9, int npan, const double* pts,
                                        int npt, double* slp, double* dlp, void* stream) {
   try {
      if (kernel != 0 && kernel != 1) throw std::runtime_error("kernel must be 0 (Laplace) or 1 (Yukawa)");
      if (npan < 0 || npt < 0) throw std::runtime_error("negative panel or point count");
      if (npan > FFTSVD_NOALLOC_MAX_PANELS)
         throw std::runtime_error("fftsvd_integrate_noalloc: " + std::to_string(npan) + " panels exceed the maximum of " +
                                  std::to_string(FFTSVD_NOALLOC_MAX_PANELS));
      if ((npan && (!verts9 || !slp || !dlp)) || (npt && !pts)) throw std::runtime_error("null array");
      if ((long long)npan * npt == 0) return 0;
      if ((long long)npan * npt >= (1LL << 31)) throw std::runtime_error("npan * npt must stay below 2^31");
      int dev = 0;
      FSVD_CHECK(hipGetDevice(&dev));
      PanelGeom* g = scratch_ptr(dev);
      hipStream_t s = (hipStream_t)stream;
      hipLaunchKernelGGL(k_panel_geometry, dim3((npan + 63) / 64), dim3(64), 0, s, verts9, npan, g);
      FSVD_CHECK(hipGetLastError());
      launch_integrate_all(kernel, kappa, g, npan, pts, npt, slp, dlp, s);
      FSVD_CHECK(hipGetLastError());
   } catch (const std::exception& e) {
      set_last_error(e.what());
      return -1;
   }
   return 0;
}
