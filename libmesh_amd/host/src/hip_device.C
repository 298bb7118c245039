// hip_device.C -- process-wide device contexts and status -> exception mapping.
#include "libmesh/hip_device.h"

#include <map>
#include <mutex>
#include <string>

namespace libMesh {

namespace {
struct Contexts {
  std::mutex m;
  std::map<int, lmh_context *> ctx;
  ~Contexts() {
    for (auto &kv : ctx) lmh_context_destroy(kv.second);
  }
};
Contexts &contexts() {
  static Contexts c;
  return c;
}
}  // namespace

lmh_context *HipDevice::context(int device) {
  Contexts &c = contexts();
  std::lock_guard<std::mutex> lock(c.m);
  auto it = c.ctx.find(device);
  if (it != c.ctx.end()) return it->second;
  lmh_context *ctx = nullptr;
  hip_check(lmh_context_create(device, &ctx), "HipDevice::context");
  c.ctx[device] = ctx;
  return ctx;
}

void hip_check(int rc, const char *what) {
  if (rc == LMH_OK) return;
  const std::string msg = std::string(what) + ": " + lmh_last_error();
  if (rc == LMH_ERR_DEGENERATE) throw DegenerateMap(msg);
  if (rc == LMH_ERR_UNSUPPORTED) throw NotImplemented();
  throw LogicError(msg);
}

}  // namespace libMesh
