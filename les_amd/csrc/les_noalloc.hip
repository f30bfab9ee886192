// les_noalloc.hip -- entry points that allocate nothing, for callers that
// cannot call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).
//
// Their scratch lives in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds from the
// module's load until the process ends.  In host code the array's name is a
// host-side stand-in, not a device address: noalloc_init looks the address
// up once per device (hipGetSymbolAddress, with that device current, from
// les_create) and only that address is handed to kernels and copies.
//
// Every call on one GPU shares that GPU's array and leaves its last contents
// behind: what must start at zero is cleared by each call, and calls that
// overlap on one GPU (other streams, other host threads) race.
#include <mutex>
#include <string>

#include "../../include/les_amd.h"
#include "les_internal.h"

static constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;  // 64 MiB

// File scope, outside every namespace: the device compiler keeps an internal
// device variable and gives it an external device symbol only when it sees
// host code take its address (noalloc_init); inside an anonymous namespace
// it drops it.  Host code elsewhere in the library uses les::scratch (the
// context arena), so this file names the array ::scratch.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on `device`, nullptr before noalloc_init
void* scratch_ptr(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace

namespace les {

int noalloc_init(int device, std::string* err) {
  if (device < 0 || device >= MAX_DEVICES) {
    if (err) *err = "noalloc: device ordinal out of range";
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p) {
    if (err) *err = std::string("noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return 1;
  }
  g_ptr[device] = p;
  return 0;
}

}  // namespace les

extern "C" {

int les_tridag_array_noalloc(les_ctx* c, const double* a, const double* b, const double* cc, const double* r,
                             double* u) {
  if (!c) return -1;
  c->err.clear();
  unsigned char* base = (unsigned char*)scratch_ptr(c->device);
  if (!base) {
    c->err = "tridag_array_noalloc: the scratch array of this device was not set up (les_create)";
    return 1;
  }
  const size_t gbytes = (les::tridag_array_gamma_bytes(c) + 255) & ~size_t(255);
  if (gbytes + 256 > NOALLOC_SCRATCH_BYTES) {
    c->err = "tridag_array_noalloc: (nz+1)*lh*ny doubles of Thomas factors exceed the " +
             std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch array; use les_tridag_array";
    return 2;
  }
  const int e = les::tridag_array_with(c, a, b, cc, r, u, (double*)base, (int*)(base + gbytes));
  if (e && c->err.empty()) c->err = e == 1 ? "tridag_array: rewrite eqs (b(1) == 0)" : "tridag_array failed (zero pivot)";
  return c->err.empty() ? 0 : 1;
}

}  // extern "C"
