// Internal C++ host layer of libheddle_hip: device context, kernel
// manifest, device buffers and the launch helper.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../kernels/hmd_params.h"
#include "heddle_hip.h"

// Kernel manifest: every entry point of the embedded code object.  The
// analogue of the reference's define_kernels! manifest
// (src/gpu/kernel_macros.rs:63-113, src/gpu/device.rs:34-52).
#define HMD_KERNEL_LIST(X)                                                   \
  X(hmd_scan_tiles) X(hmd_scan_add)                                          \
  X(hmd_nbl_displacement) X(hmd_nbl_keys) X(hmd_nbl_scatter)                 \
  X(hmd_nbl_sort_cells) X(hmd_nbl_gather) X(hmd_nbl_bbox) X(hmd_nbl_build)   \
  X(hmd_nbl_worklist) X(hmd_nbl_close) X(hmd_nbl_pairs)                     \
  X(hmd_pair_force_f) X(hmd_pair_force_fev) X(hmd_pair_force_lj_f)           \
  X(hmd_pair_force_lj_fev) X(hmd_pair_force_coul_f) X(hmd_pair_force_coul_fev)\
  X(hmd_pair_finalize) X(hmd_pair_scaled) X(hmd_slot_add)                   \
  X(hmd_spme_bin) X(hmd_spme_bin_scatter)                                    \
  X(hmd_spme_spread_p4) X(hmd_spme_gather_p4)                                \
  X(hmd_spme_spread_p5) X(hmd_spme_gather_p5)                                \
  X(hmd_spme_spread_p6) X(hmd_spme_gather_p6)                                \
  X(hmd_spme_spread_p8) X(hmd_spme_gather_p8)                                \
  X(hmd_fft_r2c_c) X(hmd_fft_c2r_c) X(hmd_fft_axis) X(hmd_spme_reduce)       \
  X(hmd_fft_axis_16) X(hmd_fft_r2c_16) X(hmd_fft_c2r_16)                     \
  X(hmd_fft_axis_32) X(hmd_fft_r2c_32) X(hmd_fft_c2r_32)                     \
  X(hmd_fft_axis_64) X(hmd_fft_r2c_64) X(hmd_fft_c2r_64)                     \
  X(hmd_fft_axis_128) X(hmd_fft_r2c_128) X(hmd_fft_c2r_128)                  \
  X(hmd_fft_axis_256) X(hmd_fft_r2c_256) X(hmd_fft_c2r_256)                  \
  X(hmd_fft_slab_fwd_128) X(hmd_fft_slab_solve_128) X(hmd_fft_slab_inv_128)   \
  X(hmd_fft_slab_fwd_64) X(hmd_fft_slab_solve_64) X(hmd_fft_slab_inv_64)      \
  X(hmd_spme_gv_table)                                                        \
  X(hmd_vv_kick_drift) X(hmd_vv_kick)                                        \
  X(hmd_settle_snapshot) X(hmd_settle_positions) X(hmd_settle_velocities)    \
  X(hmd_settle_virial)                                                       \
  X(hmd_lan_drift_half) X(hmd_lan_ou) X(hmd_counter_increment)               \
  X(hmd_step_vv_first) X(hmd_step_baoa) X(hmd_step_kick_last)                \
  X(hmd_sd_step) X(hmd_reduce_partial) X(hmd_reduce_final)                 \
  X(hmd_csvr_sample) X(hmd_csvr_finish) X(hmd_csvr_rescale) X(hmd_crescale_mu) X(hmd_crescale_positions) X(hmd_bonded_terms)                 \
  X(hmd_nn_pack) X(hmd_nn_descriptors) X(hmd_nn_desc_finalize)               \
  X(hmd_nn_forward) X(hmd_nn_backward) X(hmd_nn_grad_reduce)                 \
  X(hmd_nn_finalize) X(hmd_nn_residual) X(hmd_nn_adam) X(hmd_nn_record)      \
  X(hmd_nn_diff)

namespace hmd {

enum KernelId {
#define HMD_ENUM(name) K_##name,
  HMD_KERNEL_LIST(HMD_ENUM)
#undef HMD_ENUM
      K_COUNT
};

extern const char* const kKernelNames[K_COUNT];

// Error carrying an hmd_status; converted at the C boundary.
struct Error : std::runtime_error {
  hmd_status code;
  Error(hmd_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(HMD_ERR_GPU, std::string(what) + ": " + hipGetErrorString(e));
}

// One kernel launch as recorded instead of issued: a replica batch records
// every replica's step and merges the records into batched launches.
struct LaunchRec {
  KernelId k;
  hipFunction_t fn;        // the kernel (a composed pair kernel where one is set)
  hipFunction_t fn_batch;  // its batched twin (null: launched per replica)
  uint32_t grid, block;
  size_t shmem;
  std::vector<unsigned char> params;
};

struct Context {
  int device = 0;
  int num_cus = 0;
  size_t max_lds = 65536;   // LDS bytes one workgroup may use
  std::string device_name;
  hipStream_t stream = nullptr;
  bool owns_stream = true;   // false: the caller's stream (hmd_ctx_create_on_stream)
  hipModule_t module = nullptr;
  hipFunction_t fn[K_COUNT] = {};
  hipFunction_t fn_batch[K_COUNT] = {};   // NAME_batch twins (null where absent)
  // When set, launches are appended here instead of being issued.
  std::vector<LaunchRec>* rec = nullptr;
  // Optional per-stage event recording (profile mode, eager only).
  bool profiling = false;
  std::vector<std::pair<std::string, hipEvent_t>> marks;

  void mark(const char* stage);
};

// Typed device allocation owned by a host object.
template <class T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void alloc(size_t n, bool zero = true) {
    release();
    count = n;
    if (n == 0) return;
    check(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)), "hipMalloc");
    if (zero) {
      // hipMemset runs on the null stream, which does not order against the
      // library's non-blocking streams: wait for it, so that a kernel on
      // any stream sees the zeros.
      check(hipMemset(ptr, 0, n * sizeof(T)), "hipMemset");
      check(hipStreamSynchronize(nullptr), "sync");
    }
  }
  // Waits for the copy: the source is usually a temporary in pageable host
  // memory, which the runtime may still be reading after an asynchronous
  // copy returns.  (Never called inside a stream capture.)
  void upload(const T* host, size_t n, hipStream_t s) {
    if (n == 0) return;
    check(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s), "H2D");
    check(hipStreamSynchronize(s), "sync");
  }
  void download(T* host, size_t n, hipStream_t s) const {
    if (n == 0) return;
    check(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, s), "D2H");
    check(hipStreamSynchronize(s), "sync");
  }
};

// The launch macro of the reference (gpu_launch!, kernel_macros.rs:184-246)
// becomes one typed call: every kernel takes a single parameter struct, so
// the kernarg segment is the struct and the argument order cannot drift.
template <class P>
inline void launch(Context& c, KernelId k, uint32_t grid, uint32_t block,
                   size_t shmem, const P& params) {
  if (grid == 0) return;
  if (c.rec) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&params);
    c.rec->push_back(LaunchRec{k, c.fn[k], c.fn_batch[k], grid, block, shmem,
                               std::vector<unsigned char>(b, b + sizeof(P))});
    return;
  }
  void* args[] = {const_cast<P*>(&params)};
  check(hipModuleLaunchKernel(c.fn[k], grid, 1, 1, block, 1, 1,
                              (unsigned)shmem, c.stream, args, nullptr),
        kKernelNames[k]);
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct ForceField;

struct Particles {
  Context* ctx = nullptr;
  std::vector<ForceField*> observers;   // force fields built on these particles
  uint32_t n = 0;
  HmdLattice lattice{};
  DevBuf<float> posq;      // float4[n]
  DevBuf<float> vel;       // [3][n]
  DevBuf<float> force;     // [3][n]
  DevBuf<float> mass, inv_mass;
  DevBuf<uint32_t> type;
  DevBuf<int32_t> images;  // [3][n]
  DevBuf<double> pos_lo, vel_lo;  // [3][n] (compensated mode)
  DevBuf<float> energy, virial;   // [n]
  DevBuf<double> red_partials, red_out;
  DevBuf<unsigned long long> draw_counter;  // [1] for the standalone O step
  std::vector<float> host_mass;
  uint32_t max_type = 0;
};

// Deterministic FP64 sum (mode 0) or kinetic energy (mode 1).
double reduce_f64(Context& c, Particles& p, const float* x, uint32_t mode);

// Exclusive scan of `len` u32 (device) into out[0..len] (total at out[len]).
struct Scanner {
  DevBuf<uint32_t> tile_sums, tile_offsets;
  uint32_t len = 0;
  void init(uint32_t len_);
  void run(Context& c, const uint32_t* in, uint32_t* out, const uint32_t* gate,
           uint32_t gate_mask);
};

}  // namespace hmd

struct hmd_ctx {
  hmd::Context c;
};
struct hmd_particles {
  hmd::Particles p;
};

void hmd_set_error(const std::string& m);

// Convert C++ exceptions into an hmd_status at the C boundary.
#define HMD_GUARD(...)                                         \
  try {                                                        \
    __VA_ARGS__;                                               \
    return HMD_OK;                                             \
  } catch (const hmd::Error& e) {                              \
    hmd_set_error(e.what());                                   \
    return e.code;                                             \
  } catch (const std::exception& e) {                          \
    hmd_set_error(e.what());                                   \
    return HMD_ERR_INVALID;                                    \
  }
