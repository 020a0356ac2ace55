// Learned coarse-mesh SPME correction: the network object, its training
// data set and Adam state, and the passes that run it over a force field's
// neighbour list.  No counterpart in the reference (it has no learned
// part); the C entry points follow the hmd_mlp_{create, forward,
// backward, adam_step} boundary the hot path's C ABI reserves for it.
#include <cmath>
#include <memory>

#include "forcefield.h"

namespace hmd {

struct Mlp {
  Context& c;
  float r_cut = 0, force_scale = 1;
  uint32_t species_mask = 0;
  float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  uint32_t adam_t = 0;
  uint64_t rng = 0;
  DevBuf<float> master, adam_m, adam_v, vec;
  DevBuf<uint16_t> ops;
  DevBuf<double> grad;
  // per-system work buffers
  uint32_t n = 0, n_pad = 0, n_waves = 0;
  DevBuf<long long> desc_acc, force_acc;
  DevBuf<float> desc, desc_canon, df, resid, partials, scratch_f, target;
  DevBuf<double> red, red_out;
  uint32_t red_blocks = 0;
  // training frames
  uint32_t ds_n = 0, ds_cap = 0, ds_count = 0;
  DevBuf<float> ds_posq, ds_target;
  double ds_sum_t2 = 0.0, ds_sum_f2 = 0.0;
  std::vector<uint32_t> order;
  uint32_t order_pos = 0;

  explicit Mlp(Context& ctx) : c(ctx) {}

  uint64_t next_u64() {   // splitmix64 (Steele, Lea & Flood 2014)
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

  void pack() {
    NnPackParams pp{master.ptr, ops.ptr, vec.ptr};
    launch(c, K_hmd_nn_pack, ceil_div(3 * 4 * 2 * 64 * 8, 256), 256, 0, pp);
  }

  void init_params() {
    std::vector<float> w(HMD_NN_PARAMS);
    size_t o = 0;
    auto fill = [&](size_t count, double fan_in) {
      const double b = 1.0 / std::sqrt(fan_in);
      for (size_t k = 0; k < count; ++k) w[o++] = (float)((2.0 * uniform() - 1.0) * b);
    };
    fill(64 * 51, 51); fill(64, 51);      // W1, b1
    fill(64 * 64, 64); fill(64, 64);      // W2, b2
    fill(64, 64); fill(1, 64);            // w3, b3
    master.upload(w.data(), w.size(), c.stream);
  }

  void ensure_size(const ForceField& ff) {
    if (ff.r_search < r_cut)
      throw Error(HMD_ERR_INVALID, "network cutoff exceeds the neighbour-list search radius");
    if (ff.n == n && ff.n_pad == n_pad) return;
    n = ff.n;
    n_pad = ff.n_pad;
    desc_acc.alloc(16 * (size_t)n_pad);
    force_acc.alloc(3 * (size_t)n_pad);
    desc.alloc(16 * (size_t)n_pad);
    desc_canon.alloc(16 * (size_t)n);
    df.alloc(3 * (size_t)n);
    resid.alloc(3 * (size_t)n_pad);
    target.alloc(3 * (size_t)n);
    scratch_f.alloc(3 * (size_t)n);
    n_waves = (uint32_t)c.num_cus * 4u;
    partials.alloc((size_t)n_waves * HMD_NN_PARTIAL, false);
    red_blocks = ceil_div(n, 256);
    red.alloc(std::max(2u * 256u, red_blocks));
    red_out.alloc(4);
  }

  NnPassParams pass_params(const ForceField& ff) const {
    NnPassParams q{};
    q.posq_sorted = ff.posq_sorted.ptr;
    q.type_sorted = ff.type_sorted.ptr;
    q.work = ff.work.ptr;
    q.n_work = ff.n_work.ptr;
    q.entry_j = ff.entry_j.ptr;
    q.entry_mask = ff.entry_mask.ptr;
    q.bbox = ff.bbox.ptr;
    q.shift_margin = ff.r_cut + 0.5f * ff.r_skin;
    q.L = ff.p.lattice;
    q.n = ff.n;
    q.cap_entries = ff.cap_entries;
    q.n_pad = ff.n_pad;
    q.species_mask = species_mask;
    q.r_cut2 = r_cut * r_cut;
    q.inv_r_cut = 1.0f / r_cut;
    q.force_scale = force_scale;
    q.desc_acc = desc_acc.ptr;
    q.desc = desc.ptr;
    q.ops = ops.ptr;
    q.vec = vec.ptr;
    q.force_acc = force_acc.ptr;
    q.resid = resid.ptr;
    q.partials = partials.ptr;
    return q;
  }

  // Descriptors of the current positions (the list must be current).
  void descriptors(const ForceField& ff, bool canonical) {
    NnPassParams q = pass_params(ff);
    launch(c, K_hmd_nn_descriptors, (uint32_t)c.num_cus * 8u, 256, 0, q);
    NnDescFinalizeParams fp{desc_acc.ptr, n_pad, n, ff.sorted_id.ptr, desc.ptr,
                            canonical ? desc_canon.ptr : nullptr};
    launch(c, K_hmd_nn_desc_finalize, ceil_div(n_pad, 256), 256, 0, fp);
  }

  // Force correction into df (canonical), and added to add_to if given.
  void forward(const ForceField& ff, float* add_to) {
    NnPassParams q = pass_params(ff);
    launch(c, K_hmd_nn_forward, (uint32_t)c.num_cus * 2u, 256, 0, q);
    NnFinalizeParams fp{force_acc.ptr, n_pad, n, ff.slot_of.ptr, df.ptr, add_to};
    launch(c, K_hmd_nn_finalize, ceil_div(n, 256), 256, 0, fp);
  }

  // Residual of df against target (canonical [3][n]) for a batch of B
  // frames; the loss share goes to red_out[slot].
  void residual(const ForceField& ff, const float* tgt, uint32_t batch, uint32_t slot,
                double* loss_out) {
    const double den = 3.0 * (double)n * (double)force_scale * (double)force_scale * batch;
    NnResidualParams rp{df.ptr, tgt, ff.slot_of.ptr, n, n_pad, (float)(2.0 / den),
                        1.0 / den, resid.ptr, red.ptr};
    launch(c, K_hmd_nn_residual, red_blocks, 256, 0, rp);
    ReduceParams r{};
    r.partials = red.ptr;
    r.n_blocks = red_blocks;
    r.out = loss_out ? loss_out : red_out.ptr + slot;
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
  }

  void backward(const ForceField& ff) {
    NnPassParams q = pass_params(ff);
    launch(c, K_hmd_nn_backward, (uint32_t)c.num_cus, 256, 0, q);
    NnGradReduceParams gp{partials.ptr, n_waves, grad.ptr};
    launch(c, K_hmd_nn_grad_reduce, HMD_NN_PARAMS, 256, 0, gp);
  }

  void adam() {
    ++adam_t;
    NnAdamParams ap{master.ptr, adam_m.ptr, adam_v.ptr, grad.ptr, lr, beta1, beta2, eps,
                    (float)(1.0 - std::pow((double)beta1, (double)adam_t)),
                    (float)(1.0 - std::pow((double)beta2, (double)adam_t))};
    launch(c, K_hmd_nn_adam, ceil_div(HMD_NN_PARAMS, 256), 256, 0, ap);
    pack();
  }

  uint32_t next_frame() {   // seeded shuffle per epoch (Fisher-Yates)
    if (order.size() != ds_count || order_pos >= order.size()) {
      order.resize(ds_count);
      for (uint32_t k = 0; k < ds_count; ++k) order[k] = k;
      for (uint32_t k = ds_count; k > 1; --k) {
        const uint32_t r = (uint32_t)(next_u64() % k);
        std::swap(order[k - 1], order[r]);
      }
      order_pos = 0;
    }
    return order[order_pos++];
  }
};

// Hook of ForceField::compute: descriptors and correction of the current
// step, added to the slow-class (reciprocal) forces.
void mlp_correct(Mlp& m, ForceField& ff) {
  m.ensure_size(ff);
  m.descriptors(ff, false);
  ff.c.mark("nn_descriptors");
  m.forward(ff, ff.slow_f.ptr);
  ff.c.mark("nn_forward");
}

static void to_soa(const float* rows, uint32_t n, std::vector<float>& soa) {
  soa.resize(3 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) soa[(size_t)d * n + i] = rows[3 * (size_t)i + d];
}

static void from_soa(const std::vector<float>& soa, uint32_t n, float* rows) {
  for (uint32_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) rows[3 * (size_t)i + d] = soa[(size_t)d * n + i];
}

// An evaluation at fixed positions (a recorded frame, the judge) whose list
// overflowed has already had the capacity grown and the list rebuilt by
// check_status; nothing of it was kept, so it is simply evaluated once more.
// A second overflow is reported to the caller.
template <class Body>
static void evaluate_regrowing(Body&& body) {
  try {
    body();
  } catch (const Error& e) {
    if (e.code != HMD_ERR_NBL_OVERFLOW) throw;
    body();
  }
}

}  // namespace hmd

struct hmd_mlp {
  hmd::Mlp* m = nullptr;
};

using namespace hmd;

extern "C" {

uint32_t hmd_mlp_param_count(void) { return HMD_NN_PARAMS; }

hmd_status hmd_mlp_create(hmd_ctx* h, const hmd_mlp_desc* d, hmd_mlp** out) {
  HMD_GUARD({
    if (!h || !d || !out) throw Error(HMD_ERR_INVALID, "null argument");
    if (!(d->r_cut > 0.0) || !(d->force_scale > 0.0))
      throw Error(HMD_ERR_INVALID, "r_cut and force_scale must be positive");
    if (d->n_types == 0 || d->n_types > 32 || !d->species_of_type)
      throw Error(HMD_ERR_INVALID, "1..32 types with a species table are required");
    auto m = std::make_unique<Mlp>(h->c);
    for (uint32_t t = 0; t < d->n_types; ++t) {
      if (d->species_of_type[t] > 1u)
        throw Error(HMD_ERR_INVALID, "species must be 0 or 1");
      m->species_mask |= d->species_of_type[t] << t;
    }
    m->r_cut = (float)d->r_cut;
    m->force_scale = (float)d->force_scale;
    m->rng = d->seed;
    if (d->lr > 0.0) m->lr = (float)d->lr;
    if (d->beta1 > 0.0) m->beta1 = (float)d->beta1;
    if (d->beta2 > 0.0) m->beta2 = (float)d->beta2;
    if (d->eps > 0.0) m->eps = (float)d->eps;
    m->master.alloc(HMD_NN_PARAMS);
    m->adam_m.alloc(HMD_NN_PARAMS);
    m->adam_v.alloc(HMD_NN_PARAMS);
    m->grad.alloc(HMD_NN_PARAMS);
    m->ops.alloc(3 * 4 * 2 * 64 * 8);
    m->vec.alloc(HMD_NN_HID * 3 + 1);
    m->init_params();
    m->pack();
    check(hipStreamSynchronize(m->c.stream), "sync");
    *out = new hmd_mlp{m.release()};
  });
}

void hmd_mlp_destroy(hmd_mlp* m) {
  if (!m) return;
  delete m->m;
  delete m;
}

hmd_status hmd_mlp_set_params(hmd_mlp* w, const float* params) {
  HMD_GUARD({
    Mlp& m = *w->m;
    m.master.upload(params, HMD_NN_PARAMS, m.c.stream);
    m.pack();
    check(hipStreamSynchronize(m.c.stream), "sync");
  });
}

hmd_status hmd_mlp_get_params(hmd_mlp* w, float* params) {
  HMD_GUARD(w->m->master.download(params, HMD_NN_PARAMS, w->m->c.stream));
}

hmd_status hmd_mlp_set_force_scale(hmd_mlp* w, double s) {
  HMD_GUARD({
    if (!(s > 0.0)) throw Error(HMD_ERR_INVALID, "force_scale must be positive");
    w->m->force_scale = (float)s;
  });
}

hmd_status hmd_mlp_descriptors(hmd_mlp* w, hmd_forcefield* f, float* out) {
  HMD_GUARD({
    Mlp& m = *w->m;
    ForceField& F = *f->ff;
    m.ensure_size(F);
    F.neighbor_pre_step();
    m.descriptors(F, true);
    if (out) m.desc_canon.download(out, 16 * (size_t)F.n, m.c.stream);
    check(hipStreamSynchronize(m.c.stream), "sync");
    F.check_status();
  });
}

hmd_status hmd_mlp_forward(hmd_mlp* w, hmd_forcefield* f, float* dforce) {
  HMD_GUARD({
    Mlp& m = *w->m;
    ForceField& F = *f->ff;
    m.ensure_size(F);
    F.neighbor_pre_step();
    m.descriptors(F, false);
    m.forward(F, nullptr);
    check(hipStreamSynchronize(m.c.stream), "sync");
    F.check_status();
    if (dforce) {
      std::vector<float> soa(3 * (size_t)F.n);
      m.df.download(soa.data(), soa.size(), m.c.stream);
      from_soa(soa, F.n, dforce);
    }
  });
}

hmd_status hmd_mlp_loss_grad(hmd_mlp* w, hmd_forcefield* f, const float* target,
                             double* loss, float* grad) {
  HMD_GUARD({
    Mlp& m = *w->m;
    ForceField& F = *f->ff;
    m.ensure_size(F);
    std::vector<float> soa;
    to_soa(target, F.n, soa);
    m.target.upload(soa.data(), soa.size(), m.c.stream);
    check(hipMemsetAsync(m.grad.ptr, 0, HMD_NN_PARAMS * sizeof(double), m.c.stream), "memset");
    F.neighbor_pre_step();
    m.descriptors(F, false);
    m.forward(F, nullptr);
    m.residual(F, m.target.ptr, 1, 0, nullptr);
    m.backward(F);
    double l = 0.0;
    m.red_out.download(&l, 1, m.c.stream);
    F.check_status();
    if (loss) *loss = l;
    std::vector<double> g(HMD_NN_PARAMS);
    m.grad.download(g.data(), g.size(), m.c.stream);
    check(hipMemsetAsync(m.grad.ptr, 0, HMD_NN_PARAMS * sizeof(double), m.c.stream), "memset");
    if (grad)
      for (uint32_t k = 0; k < HMD_NN_PARAMS; ++k) grad[k] = (float)g[k];
  });
}

hmd_status hmd_mlp_adam_step(hmd_mlp* w, const float* grad) {
  HMD_GUARD({
    Mlp& m = *w->m;
    std::vector<double> g(grad, grad + HMD_NN_PARAMS);
    m.grad.upload(g.data(), g.size(), m.c.stream);
    m.adam();
    check(hipStreamSynchronize(m.c.stream), "sync");
  });
}

hmd_status hmd_mlp_attach(hmd_forcefield* f, hmd_mlp* w) {
  HMD_GUARD({
    ForceField& F = *f->ff;
    if (w) {
      if (!F.spme_on)
        throw Error(HMD_ERR_INVALID, "the correction needs an spme-reciprocal slot");
      w->m->ensure_size(F);
    }
    F.mlp = w ? w->m : nullptr;
    F.graph_invalid = true;
  });
}

hmd_status hmd_mlp_dataset_reserve(hmd_mlp* w, uint32_t n_atoms, uint32_t n_frames) {
  HMD_GUARD({
    Mlp& m = *w->m;
    m.ds_n = n_atoms;
    m.ds_cap = n_frames;
    m.ds_count = 0;
    m.ds_sum_t2 = m.ds_sum_f2 = 0.0;
    m.order.clear();
    m.ds_posq.alloc(4 * (size_t)n_atoms * n_frames, false);
    m.ds_target.alloc(3 * (size_t)n_atoms * n_frames, false);
  });
}

hmd_status hmd_mlp_record_frame(hmd_mlp* w, hmd_forcefield* fine, hmd_forcefield* coarse) {
  HMD_GUARD({
    Mlp& m = *w->m;
    ForceField& A = *fine->ff;
    ForceField& B = *coarse->ff;
    if (&A.p != &B.p) throw Error(HMD_ERR_INVALID, "fine and coarse must share particles");
    if (!A.spme_on || !B.spme_on) throw Error(HMD_ERR_INVALID, "both need spme-reciprocal");
    if (A.n != m.ds_n) throw Error(HMD_ERR_INVALID, "data set reserved for another size");
    if (m.ds_count >= m.ds_cap) throw Error(HMD_ERR_INVALID, "data set is full");
    m.ensure_size(B);
    evaluate_regrowing([&] {
    A.neighbor_pre_step();
    A.spme_forces(false);
    B.neighbor_pre_step();
    B.spme_forces(false);
    const size_t f = m.ds_count;
    NnRecordParams rp{A.p.posq.ptr, A.slow_f.ptr, B.slow_f.ptr, A.n,
                      m.ds_posq.ptr + 4 * (size_t)A.n * f, m.ds_target.ptr + 3 * (size_t)A.n * f};
    launch(m.c, K_hmd_nn_record, ceil_div(A.n, 256), 256, 0, rp);
    NnDiffParams dp{A.slow_f.ptr, B.slow_f.ptr, A.n, m.red.ptr};
    launch(m.c, K_hmd_nn_diff, 256, 256, 0, dp);
    ReduceParams r{};
    r.partials = m.red.ptr;
    r.n_blocks = 256;
    r.out = m.red_out.ptr + 2;
    launch(m.c, K_hmd_reduce_final, 1, 256, 0, r);
    r.partials = m.red.ptr + 256;
    r.out = m.red_out.ptr + 3;
    launch(m.c, K_hmd_reduce_final, 1, 256, 0, r);
    double s[2];
    check(hipMemcpyAsync(s, m.red_out.ptr + 2, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         m.c.stream), "D2H");
    check(hipStreamSynchronize(m.c.stream), "sync");
    A.check_status();
    B.check_status();
    m.ds_sum_t2 += s[0];
    m.ds_sum_f2 += s[1];
    ++m.ds_count;
    });
  });
}

hmd_status hmd_mlp_dataset_stats(hmd_mlp* w, double out[3]) {
  HMD_GUARD({
    Mlp& m = *w->m;
    const double comps = 3.0 * (double)m.ds_n * (double)std::max(1u, m.ds_count);
    out[0] = (double)m.ds_count;
    out[1] = std::sqrt(m.ds_sum_t2 / comps);
    out[2] = std::sqrt(m.ds_sum_f2 / comps);
  });
}

hmd_status hmd_mlp_train(hmd_mlp* w, hmd_forcefield* f, uint32_t n_steps, uint32_t batch,
                         double* losses) {
  HMD_GUARD({
    Mlp& m = *w->m;
    ForceField& F = *f->ff;
    if (m.ds_count == 0) throw Error(HMD_ERR_INVALID, "no recorded frames");
    if (F.n != m.ds_n) throw Error(HMD_ERR_INVALID, "force field size differs from the data");
    if (batch == 0) throw Error(HMD_ERR_INVALID, "batch must be positive");
    m.ensure_size(F);
    DevBuf<double> frame_loss;
    frame_loss.alloc((size_t)n_steps * batch);
    check(hipMemsetAsync(m.grad.ptr, 0, HMD_NN_PARAMS * sizeof(double), m.c.stream), "memset");
    for (uint32_t s = 0; s < n_steps; ++s) {
      for (uint32_t b = 0; b < batch; ++b) {
        const size_t fr = m.next_frame();
        check(hipMemcpyAsync(F.p.posq.ptr, m.ds_posq.ptr + 4 * (size_t)F.n * fr,
                             4 * (size_t)F.n * sizeof(float), hipMemcpyDeviceToDevice,
                             m.c.stream), "D2D");
        F.neighbor_pre_step();
        m.descriptors(F, false);
        m.forward(F, nullptr);
        m.residual(F, m.ds_target.ptr + 3 * (size_t)F.n * fr, batch, 0,
                   frame_loss.ptr + (size_t)s * batch + b);
        m.backward(F);
      }
      m.adam();
      // an overflowing list would silently drop pairs: check every step
      check(hipStreamSynchronize(m.c.stream), "sync");
      F.check_status();
    }
    std::vector<double> fl((size_t)n_steps * batch);
    frame_loss.download(fl.data(), fl.size(), m.c.stream);
    if (losses)
      for (uint32_t s = 0; s < n_steps; ++s) {
        double t = 0.0;
        for (uint32_t b = 0; b < batch; ++b) t += fl[(size_t)s * batch + b];
        losses[s] = t;
      }
  });
}

hmd_status hmd_mlp_judge(hmd_forcefield* nn, hmd_forcefield* fine, double* rel_rms) {
  HMD_GUARD({
    ForceField& A = *nn->ff;
    ForceField& B = *fine->ff;
    if (&A.p != &B.p) throw Error(HMD_ERR_INVALID, "both force fields must share particles");
    if (!A.mlp) throw Error(HMD_ERR_INVALID, "no network attached to the first force field");
    Mlp& m = *A.mlp;
    Context& c = A.c;
    evaluate_regrowing([&] {
    B.compute(0);
    check(hipMemcpyAsync(m.scratch_f.ptr, B.p.force.ptr, 3 * (size_t)B.n * sizeof(float),
                         hipMemcpyDeviceToDevice, c.stream), "D2D");
    A.compute(0);
    NnDiffParams dp{A.p.force.ptr, m.scratch_f.ptr, A.n, m.red.ptr};
    launch(c, K_hmd_nn_diff, 256, 256, 0, dp);
    ReduceParams r{};
    r.partials = m.red.ptr;
    r.n_blocks = 256;
    r.out = m.red_out.ptr + 2;
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
    r.partials = m.red.ptr + 256;
    r.out = m.red_out.ptr + 3;
    launch(c, K_hmd_reduce_final, 1, 256, 0, r);
    double s[2];
    check(hipMemcpyAsync(s, m.red_out.ptr + 2, 2 * sizeof(double), hipMemcpyDeviceToHost,
                         c.stream), "D2H");
    check(hipStreamSynchronize(c.stream), "sync");
    A.check_status();
    B.check_status();
    *rel_rms = s[1] > 0.0 ? std::sqrt(s[0] / s[1]) : 0.0;
    });
  });
}

}  // extern "C"
