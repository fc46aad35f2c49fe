// noalloc.hip -- the element coefficient gradient without device allocations.
//
// bddcml_amd_element_gradient allocates its device buffers on every call
// (element dof table, element ids, element matrices, u, lam, gradient).  The
// variant here keeps them in one file-scope device array, `scratch`, of
// kScratchBytes, so a caller that must not allocate device memory (a tight
// loop of gradient evaluations, a harness that forbids hipMalloc) can call
// it.  Every GPU the library runs on holds its own copy of the array until
// the process ends; its address on the current device is taken once with
// hipGetSymbolAddress (the array's name in host code is not a device
// address) and only that address is handed to kernels and copies.  Calls on
// one GPU share the array: calls that overlap on one GPU race, and each call
// clears the part it accumulates into.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "solver.hpp"

namespace bddc {

#define NA_CHECK(x)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      fatal("BDDCML_AMD_ELEMENT_GRADIENT_NOALLOC", std::string(#x) + " failed: " +       \
                                                       hipGetErrorString(e_));            \
  } while (0)

// 64 MiB: a 64^3-element Q1 problem needs about 16 MiB
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch array, looked up on first use
static unsigned char *scratch_on_current_device(const char *R) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) fatal(R, "no current device");
  if (dev < 0 || dev >= kMaxDevices) fatal(R, "device index beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fatal(R, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (unsigned char *)p;
  }
  return g_scratch_ptr[dev];
}

__global__ void k_fill_zero(size_t n, double *__restrict__ v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// compute stream of a set-up solver (device.hip)
hipStream_t device_stream(Solver &S);

void device_element_gradient_noalloc(Solver &S, const double *ae, int nke, bool per_element,
                                     const double *u, const double *lam, double *grad) {
  const char *R = "BDDCML_AMD_ELEMENT_GRADIENT_NOALLOC";
  hipStream_t st = device_stream(S);
  // element dof lists in global numbering and global element ids (host)
  std::vector<int> edof, eid;
  for (const Sub &s : S.subs) {
    if ((int)s.isegn.size() != s.nelems) fatal(R, "global element numbers (isegn) were not uploaded");
    int k = 0;
    for (int e = 0; e < s.nelems; e++) {
      int nd = 0;
      for (int q = 0; q < s.nnet[e]; q++) {
        const int l = s.inet[k + q];
        for (int d = s.kdof[l]; d < s.kdof[l + 1]; d++, nd++) edof.push_back(s.isvgvn[d]);
      }
      k += s.nnet[e];
      if (nd != nke) fatal(R, "element dof count differs from nke");
      if (s.isegn[e] < 0 || s.isegn[e] >= S.nelem_g) fatal(R, "element number out of range");
      eid.push_back(s.isegn[e]);
    }
  }
  for (int d : edof)
    if (d < 0 || d >= S.ndof_g) fatal(R, "dof number out of range");
  const int nel = (int)eid.size();
  const size_t nae = per_element ? (size_t)S.nelem_g * nke * nke : (size_t)nke * nke;
  // layout in the scratch array, each part 256-byte aligned
  const size_t b_edof = align256(edof.size() * sizeof(int)), b_eid = align256(eid.size() * sizeof(int)),
               b_ae = align256(nae * sizeof(double)), b_vec = align256((size_t)S.ndof_g * sizeof(double)),
               b_grad = align256((size_t)S.nelem_g * sizeof(double));
  const size_t need = b_edof + b_eid + b_ae + 2 * b_vec + b_grad;
  if (need > kScratchBytes)
    fatal(R, "problem needs " + std::to_string(need) + " bytes of scratch, more than the " +
                 std::to_string(kScratchBytes) + " bytes of this entry point; use "
                 "bddcml_amd_element_gradient");
  unsigned char *base = scratch_on_current_device(R);
  int *dedof = (int *)base;
  int *deid = (int *)(base + b_edof);
  double *dae = (double *)(base + b_edof + b_eid);
  double *du = (double *)(base + b_edof + b_eid + b_ae);
  double *dl = (double *)(base + b_edof + b_eid + b_ae + b_vec);
  double *dg = (double *)(base + b_edof + b_eid + b_ae + 2 * b_vec);
  NA_CHECK(hipMemcpyAsync(dedof, edof.data(), edof.size() * sizeof(int), hipMemcpyHostToDevice, st));
  NA_CHECK(hipMemcpyAsync(deid, eid.data(), eid.size() * sizeof(int), hipMemcpyHostToDevice, st));
  NA_CHECK(hipMemcpyAsync(dae, ae, nae * sizeof(double), hipMemcpyHostToDevice, st));
  NA_CHECK(hipMemcpyAsync(du, u, (size_t)S.ndof_g * sizeof(double), hipMemcpyHostToDevice, st));
  NA_CHECK(hipMemcpyAsync(dl, lam, (size_t)S.ndof_g * sizeof(double), hipMemcpyHostToDevice, st));
  // the gradient entries of elements no subdomain holds stay zero: cleared
  // by a kernel (the array keeps the previous call's values)
  const size_t ng = (size_t)S.nelem_g;
  if (ng > 0)
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, ng, dg);
  launch_element_gradient(st, nel, nke, dedof, deid, dae, per_element, du, dl, dg);
  NA_CHECK(hipGetLastError());
  NA_CHECK(hipMemcpyAsync(grad, dg, ng * sizeof(double), hipMemcpyDeviceToHost, st));
  NA_CHECK(hipStreamSynchronize(st));
}

}  // namespace bddc
