// Device context, code-object loading, particle buffers, scans and FP64
// reductions of libheddle_hip.
#include <algorithm>
#include <cstdio>

#include <cmath>

#include "forcefield.h"
#include "hmd_internal.h"

// The gfx950 code object, embedded at link time (kernels_blob.S).
extern "C" const unsigned char hmd_code_object[];

namespace hmd {

const char* const kKernelNames[K_COUNT] = {
#define HMD_NAME(name) #name,
    HMD_KERNEL_LIST(HMD_NAME)
#undef HMD_NAME
};

void Context::mark(const char* stage) {
  if (!profiling) return;
  hipEvent_t e;
  check(hipEventCreate(&e), "hipEventCreate");
  check(hipEventRecord(e, stream), "hipEventRecord");
  marks.emplace_back(stage, e);
}

void Scanner::init(uint32_t len_) {
  len = len_;
  uint32_t tiles = ceil_div(len + 1, 4096);
  if (tiles > 4096) throw Error(HMD_ERR_INVALID, "scan length exceeds 16.7M");
  tile_sums.alloc(tiles);
  tile_offsets.alloc(tiles + 1);
}

// Exclusive scan of in[0..len) into out[0..len] (out[len] = total).
void Scanner::run(Context& c, const uint32_t* in, uint32_t* out,
                  const uint32_t* gate, uint32_t gate_mask) {
  uint32_t tiles = ceil_div(len, 4096);
  if (tiles <= 1) {
    ScanParams p{in, out, nullptr, len, gate, gate_mask, 1u};
    launch(c, K_hmd_scan_tiles, 1, 1024, 0, p);
    return;
  }
  ScanParams p1{in, out, tile_sums.ptr, len, gate, gate_mask, 0u};
  launch(c, K_hmd_scan_tiles, tiles, 1024, 0, p1);
  ScanParams p2{tile_sums.ptr, tile_offsets.ptr, nullptr, tiles, gate, gate_mask, 1u};
  launch(c, K_hmd_scan_tiles, 1, 1024, 0, p2);
  ScanParams p3{tile_offsets.ptr, out, tile_sums.ptr, len, gate, gate_mask, 1u};
  launch(c, K_hmd_scan_add, tiles, 1024, 0, p3);
}

double reduce_f64(Context& c, Particles& p, const float* x, uint32_t mode) {
  const uint32_t blocks = p.red_partials.count;
  ReduceParams r{};
  r.x = x;
  r.v = p.vel.ptr;
  r.m = p.mass.ptr;
  r.n = p.n;
  r.mode = mode;
  r.partials = p.red_partials.ptr;
  r.out = p.red_out.ptr;
  r.n_blocks = blocks;
  launch(c, K_hmd_reduce_partial, blocks, 256, 0, r);
  launch(c, K_hmd_reduce_final, 1, 256, 0, r);
  double out = 0.0;
  p.red_out.download(&out, 1, c.stream);
  return out;
}

}  // namespace hmd

using namespace hmd;

static thread_local std::string g_last_error;
void hmd_set_error(const std::string& m) { g_last_error = m; }

extern "C" {

const char* hmd_last_error(void) { return g_last_error.c_str(); }
const char* hmd_version(void) { return "heddle_hip 0.1 (gfx950)"; }

int hmd_kernel_count(void) { return K_COUNT; }
const char* hmd_kernel_name(int i) {
  return (i >= 0 && i < K_COUNT) ? kKernelNames[i] : nullptr;
}

static hmd_status ctx_create(int device, hipStream_t caller_stream, hmd_ctx** out) {
  HMD_GUARD({
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device)
      throw Error(HMD_ERR_NO_DEVICE, "no HIP device " + std::to_string(device));
    auto* h = new hmd_ctx();
    try {
      Context& c = h->c;
      c.device = device;
      check(hipSetDevice(device), "hipSetDevice");
      hipDeviceProp_t prop;
      check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
      c.num_cus = prop.multiProcessorCount;
      c.max_lds = prop.sharedMemPerBlock;
      c.device_name = prop.gcnArchName;
      if (c.device_name.rfind("gfx950", 0) != 0)
        throw Error(HMD_ERR_UNSUPPORTED,
                    "code object is gfx950-only; device is " + c.device_name);
      if (caller_stream) {
        c.stream = caller_stream;
        c.owns_stream = false;
      } else {
        check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
      }
      check(hipModuleLoadData(&c.module, hmd_code_object), "hipModuleLoadData");
      for (int k = 0; k < K_COUNT; ++k) {
        check(hipModuleGetFunction(&c.fn[k], c.module, kKernelNames[k]), kKernelNames[k]);
        const std::string bn = std::string(kKernelNames[k]) + "_batch";
        if (hipModuleGetFunction(&c.fn_batch[k], c.module, bn.c_str()) != hipSuccess) {
          c.fn_batch[k] = nullptr;
          (void)hipGetLastError();
        }
      }
    } catch (...) {
      hmd_ctx_destroy(h);
      throw;
    }
    *out = h;
  });
}

hmd_status hmd_ctx_create(int device, hmd_ctx** out) { return ctx_create(device, nullptr, out); }

hmd_status hmd_ctx_create_on_stream(int device, void* stream, hmd_ctx** out) {
  if (!stream) {
    hmd_set_error("null stream");
    return HMD_ERR_INVALID;
  }
  return ctx_create(device, static_cast<hipStream_t>(stream), out);
}

void hmd_ctx_destroy(hmd_ctx* h) {
  if (!h) return;
  Context& c = h->c;
  for (auto& m : c.marks) (void)hipEventDestroy(m.second);
  if (c.module) (void)hipModuleUnload(c.module);
  if (c.stream && c.owns_stream) (void)hipStreamDestroy(c.stream);
  delete h;
}

hmd_status hmd_ctx_synchronize(hmd_ctx* h) {
  HMD_GUARD(check(hipStreamSynchronize(h->c.stream), "hipStreamSynchronize"));
}

hmd_status hmd_ctx_device_info(hmd_ctx* h, int* cus, char* name, int len) {
  HMD_GUARD({
    if (cus) *cus = h->c.num_cus;
    if (name && len > 0) std::snprintf(name, (size_t)len, "%s", h->c.device_name.c_str());
  });
}

hmd_status hmd_particles_create(hmd_ctx* h, uint32_t n, const float* posq,
                                const float* vel, const float* mass,
                                const uint32_t* type, const float lat[6],
                                hmd_particles** out) {
  HMD_GUARD({
    if ((n && (!posq || !mass)) || !lat) throw Error(HMD_ERR_INVALID, "particles: null input");
    for (uint32_t i = 0; i < n; ++i)
      if (!(mass[i] > 0.0f)) throw Error(HMD_ERR_INVALID, "particles: non-positive mass");
    auto* hp = new hmd_particles();
    Particles& p = hp->p;
    p.ctx = &h->c;
    p.n = n;
    p.lattice = HmdLattice{lat[0], lat[1], lat[2], lat[3], lat[4], lat[5]};
    hipStream_t s = h->c.stream;
    p.posq.alloc(4 * (size_t)n);
    p.vel.alloc(3 * (size_t)n);
    p.force.alloc(3 * (size_t)n);
    p.mass.alloc(n);
    p.inv_mass.alloc(n);
    p.type.alloc(n);
    p.images.alloc(3 * (size_t)n);
    p.pos_lo.alloc(3 * (size_t)n);
    p.vel_lo.alloc(3 * (size_t)n);
    p.energy.alloc(n);
    p.virial.alloc(n);
    p.red_partials.alloc(std::min<uint32_t>(1024u, ceil_div(n, 256)));
    p.red_out.alloc(1);
    p.draw_counter.alloc(1);
    p.host_mass.assign(mass, mass + n);
    std::vector<float> inv(n);
    for (uint32_t i = 0; i < n; ++i) inv[i] = 1.0f / mass[i];
    p.posq.upload(posq, 4 * (size_t)n, s);
    p.mass.upload(mass, n, s);
    p.inv_mass.upload(inv.data(), n, s);
    std::vector<uint32_t> ty(n, 0u);
    if (type) ty.assign(type, type + n);
    for (uint32_t t : ty) p.max_type = std::max(p.max_type, t);
    p.type.upload(ty.data(), n, s);
    if (vel) {
      std::vector<float> soa(3 * (size_t)n);
      for (uint32_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) soa[(size_t)d * n + i] = vel[3 * (size_t)i + d];
      p.vel.upload(soa.data(), soa.size(), s);
    }
    check(hipStreamSynchronize(s), "sync");
    *out = hp;
  });
}

void hmd_particles_destroy(hmd_particles* p) { delete p; }

static void soa_to_aos(const std::vector<float>& soa, float* aos, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) aos[3 * (size_t)i + d] = soa[(size_t)d * n + i];
}

hmd_status hmd_particles_download(hmd_particles* hp, float* posq, float* vel,
                                  float* force, int32_t* images) {
  HMD_GUARD({
    Particles& p = hp->p;
    hipStream_t s = p.ctx->stream;
    if (posq) p.posq.download(posq, 4 * (size_t)p.n, s);
    std::vector<float> soa(3 * (size_t)p.n);
    if (vel) { p.vel.download(soa.data(), soa.size(), s); soa_to_aos(soa, vel, p.n); }
    if (force) { p.force.download(soa.data(), soa.size(), s); soa_to_aos(soa, force, p.n); }
    if (images) {
      std::vector<int32_t> im(3 * (size_t)p.n);
      p.images.download(im.data(), im.size(), s);
      for (uint32_t i = 0; i < p.n; ++i)
        for (int d = 0; d < 3; ++d) images[3 * (size_t)i + d] = im[(size_t)d * p.n + i];
    }
  });
}

hmd_status hmd_particles_upload(hmd_particles* hp, const float* posq, const float* vel) {
  HMD_GUARD({
    Particles& p = hp->p;
    hipStream_t s = p.ctx->stream;
    // the uploaded f32 values are the whole new state: the compensated low
    // parts of what is replaced start again at zero
    if (posq) {
      p.posq.upload(posq, 4 * (size_t)p.n, s);
      if (p.pos_lo.ptr && p.pos_lo.count)
        check(hipMemsetAsync(p.pos_lo.ptr, 0, p.pos_lo.count * sizeof(double), s), "memset");
    }
    if (vel) {
      std::vector<float> soa(3 * (size_t)p.n);
      for (uint32_t i = 0; i < p.n; ++i)
        for (int d = 0; d < 3; ++d) soa[(size_t)d * p.n + i] = vel[3 * (size_t)i + d];
      p.vel.upload(soa.data(), soa.size(), s);
      if (p.vel_lo.ptr && p.vel_lo.count)
        check(hipMemsetAsync(p.vel_lo.ptr, 0, p.vel_lo.count * sizeof(double), s), "memset");
    }
    check(hipStreamSynchronize(s), "sync");
  });
}

hmd_status hmd_particles_set_box(hmd_particles* hp, const float lat[6]) {
  HMD_GUARD({
    if (!hp || !lat) throw Error(HMD_ERR_INVALID, "null argument");
    for (int k = 0; k < 6; ++k)
      if (!std::isfinite(lat[k])) throw Error(HMD_ERR_INVALID, "lattice is not finite");
    if (!(lat[0] > 0.0f && lat[1] > 0.0f && lat[2] > 0.0f))
      throw Error(HMD_ERR_INVALID, "lattice lengths must be positive");
    Particles& p = hp->p;
    const HmdLattice old = p.lattice;
    p.lattice = HmdLattice{lat[0], lat[1], lat[2], lat[3], lat[4], lat[5]};
    try {
      for (ForceField* f : p.observers) f->on_box_change();
    } catch (...) {
      p.lattice = old;   // a box the force fields refuse leaves the old one in place
      for (ForceField* f : p.observers) {
        try { f->on_box_change(); } catch (...) {}
      }
      throw;
    }
  });
}

hmd_status hmd_particles_get_box(hmd_particles* hp, float lat[6]) {
  HMD_GUARD({
    if (!hp || !lat) throw Error(HMD_ERR_INVALID, "null argument");
    const HmdLattice& L = hp->p.lattice;
    lat[0] = L.lx; lat[1] = L.ly; lat[2] = L.lz;
    lat[3] = L.xy; lat[4] = L.xz; lat[5] = L.yz;
  });
}

hmd_status hmd_particles_upload_images(hmd_particles* hp, const int32_t* images) {
  HMD_GUARD({
    if (!hp || !images) throw Error(HMD_ERR_INVALID, "null argument");
    Particles& p = hp->p;
    std::vector<int32_t> soa(3 * (size_t)p.n);
    for (uint32_t i = 0; i < p.n; ++i)
      for (int d = 0; d < 3; ++d) soa[(size_t)d * p.n + i] = images[3 * (size_t)i + d];
    p.images.upload(soa.data(), soa.size(), p.ctx->stream);
    check(hipStreamSynchronize(p.ctx->stream), "sync");
  });
}

hmd_status hmd_particles_upload_forces(hmd_particles* hp, const float* force) {
  HMD_GUARD({
    if (!hp || !force) throw Error(HMD_ERR_INVALID, "null argument");
    Particles& p = hp->p;
    std::vector<float> soa(3 * (size_t)p.n);
    for (uint32_t i = 0; i < p.n; ++i)
      for (int d = 0; d < 3; ++d) soa[(size_t)d * p.n + i] = force[3 * (size_t)i + d];
    p.force.upload(soa.data(), soa.size(), p.ctx->stream);
    check(hipStreamSynchronize(p.ctx->stream), "sync");
  });
}

hmd_status hmd_particles_negate_velocities(hmd_particles* hp) {
  HMD_GUARD({
    Particles& p = hp->p;
    hipStream_t s = p.ctx->stream;
    std::vector<float> v(3 * (size_t)p.n);
    p.vel.download(v.data(), v.size(), s);
    for (float& x : v) x = -x;
    p.vel.upload(v.data(), v.size(), s);
    if (p.vel_lo.ptr) {
      std::vector<double> lo(3 * (size_t)p.n);
      p.vel_lo.download(lo.data(), lo.size(), s);
      for (double& x : lo) x = -x;
      p.vel_lo.upload(lo.data(), lo.size(), s);
    }
    check(hipStreamSynchronize(s), "sync");
  });
}

hmd_status hmd_particles_download_energies(hmd_particles* hp, float* e, float* w) {
  HMD_GUARD({
    Particles& p = hp->p;
    if (e) p.energy.download(e, p.n, p.ctx->stream);
    if (w) p.virial.download(w, p.n, p.ctx->stream);
  });
}

hmd_status hmd_particles_download_lo(hmd_particles* hp, double* pos_lo, double* vel_lo) {
  HMD_GUARD({
    Particles& p = hp->p;
    std::vector<double> soa(3 * (size_t)p.n);
    auto tr = [&](double* out) {
      for (uint32_t i = 0; i < p.n; ++i)
        for (int d = 0; d < 3; ++d) out[3 * (size_t)i + d] = soa[(size_t)d * p.n + i];
    };
    if (pos_lo) { p.pos_lo.download(soa.data(), soa.size(), p.ctx->stream); tr(pos_lo); }
    if (vel_lo) { p.vel_lo.download(soa.data(), soa.size(), p.ctx->stream); tr(vel_lo); }
  });
}

hmd_status hmd_particles_kinetic_energy(hmd_particles* hp, double* ke) {
  HMD_GUARD(*ke = reduce_f64(*hp->p.ctx, hp->p, nullptr, 1));
}

hmd_status hmd_particles_energy_sums(hmd_particles* hp, double* pe, double* vir) {
  HMD_GUARD({
    if (pe) *pe = reduce_f64(*hp->p.ctx, hp->p, hp->p.energy.ptr, 0);
    if (vir) *vir = reduce_f64(*hp->p.ctx, hp->p, hp->p.virial.ptr, 0);
  });
}

}  // extern "C"
