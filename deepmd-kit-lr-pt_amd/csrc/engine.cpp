// Whole-path se_e2_a engine: packed frames -> E, F, V on one GPU.
//
// Host orchestration of the HIP kernels for the path of the reference TF
// descriptor (deepmd/tf/descriptor/se_a.py:578-779) and PT model
// (deepmd/pt/model/descriptor/se_a.py:742-859, task/fitting.py:1079-1104,
// model/transform_output.py:70-211): cell list -> neighbour list ->
// environment matrix -> table embedding + contraction -> descriptor ->
// fitting (per-type row blocks) -> backward to net_deriv -> force/virial ->
// ghost fold and per-frame reductions.  Everything after the per-step box
// upload runs asynchronously on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "dpmi_common.h"
#include "engine.h"
#include "kernels.h"

namespace dpmi {

namespace {

template <typename T>
T* dalloc(size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  DPMI_CHECK(hipMalloc(&p, n * sizeof(T)));
  return static_cast<T*>(p);
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
T* dupload(const T* h, size_t n) {
  T* d = dalloc<T>(n);
  if (n) DPMI_CHECK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}


// (XCD count, compute units) of the current device
std::pair<int, int> chip_layout() {
  int dev = 0, n = 1, cus = 1;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess)
    n = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 1;
  return {std::max(n, 1), std::max(cus, 1)};
}

}  // namespace

// Row tiles of the fused fitting kernel, (type, first row, rows, 0), in launch
// order.  Workgroups are dealt round-robin to the nx XCDs (block b runs on
// XCD b % nx), each with its own L2: when a launch fits in one pass, the XCDs
// are split between types in proportion to their tile counts so that each L2
// holds the weights of one type; padding slots carry zero rows.
std::vector<int4> fitting_tile_order(const std::vector<int>& tstart, int ntypes, int rows_per_tile, int nx,
                                     int cus) {
  std::vector<std::vector<int4>> per(ntypes);
  int total = 0, used = 0;
  for (int t = 0; t < ntypes; ++t) {
    for (int r = tstart[t]; r < tstart[t + 1]; r += rows_per_tile)
      per[t].push_back(make_int4(t, r, std::min(rows_per_tile, tstart[t + 1] - r), 0));
    total += (int)per[t].size();
    used += per[t].empty() ? 0 : 1;
  }
  std::vector<int4> plain;
  for (auto& v : per) plain.insert(plain.end(), v.begin(), v.end());
  if (used <= 1 || used > nx || nx <= 1) return plain;
  // XCDs per type: largest remainder of the proportional share, at least one
  std::vector<int> nxt(ntypes, 0);
  int given = 0;
  for (int t = 0; t < ntypes; ++t)
    if (!per[t].empty()) {
      nxt[t] = std::max(1, (int)((long long)nx * per[t].size() / total));
      given += nxt[t];
    }
  while (given > nx) {
    int best = -1;
    for (int t = 0; t < ntypes; ++t)
      if (nxt[t] > 1 && (best < 0 || per[t].size() * nxt[best] < per[best].size() * nxt[t])) best = t;
    --nxt[best];
    --given;
  }
  while (given < nx) {  // give the next XCD to the type with the most tiles per XCD
    int best = -1;
    for (int t = 0; t < ntypes; ++t)
      if (nxt[t] > 0 && (best < 0 || per[t].size() * nxt[best] > per[best].size() * nxt[t])) best = t;
    ++nxt[best];
    ++given;
  }
  std::vector<std::vector<int4>> lanes(nx);
  int x = 0;
  for (int t = 0; t < ntypes; ++t) {
    for (size_t k = 0; k < per[t].size(); ++k) lanes[x + (int)(k % nxt[t])].push_back(per[t][k]);
    x += nxt[t];
  }
  size_t depth = 0;
  for (auto& l : lanes) depth = std::max(depth, l.size());
  std::vector<int4> out(depth * nx, make_int4(0, 0, 0, 0));
  for (int xi = 0; xi < nx; ++xi)
    for (size_t k = 0; k < lanes[xi].size(); ++k) out[k * nx + xi] = lanes[xi][k];
  // keep the padded order only when it needs no more passes over the
  // chip (one workgroup per CU) than the plain one
  const long long per_xcd = std::max(1, cus / nx);
  const long long passes_out = ((long long)depth + per_xcd - 1) / per_xcd;
  const long long passes_plain = ((long long)(plain.size() + nx - 1) / nx + per_xcd - 1) / per_xcd;
  return passes_out <= passes_plain ? out : plain;
}

// ----------------------------------------------------------------- model --
SeaModel::SeaModel(const dpmi_sea_model_desc& d) {
  ntypes = d.ntypes;
  if (ntypes <= 0 || ntypes > 127) throw std::invalid_argument("ntypes must be in [1, 127]");
  sel.assign(d.sel, d.sel + ntypes);
  sec.assign(ntypes + 1, 0);
  for (int t = 0; t < ntypes; ++t) sec[t + 1] = sec[t] + sel[t];
  nnei = sec[ntypes];
  rcut = d.rcut;
  rcut_smth = d.rcut_smth;
  M = d.last_layer_size;
  A = d.axis_neuron;
  if (A > M) throw std::invalid_argument("axis_neuron must not exceed the last embedding width");
  if (A > 16) throw std::invalid_argument("axis_neuron above 16 is not supported by the fused descriptor kernels");
  if (M > 1024) throw std::invalid_argument("last_layer_size must be at most 1024");
  type_one_side = d.type_one_side != 0;
  nnets = d.nnets;
  if (nnets != (type_one_side ? ntypes : ntypes * ntypes)) throw std::invalid_argument("nnets mismatch");
  davg.assign(d.davg, d.davg + (size_t)ntypes * nnei * 4);
  dstd.assign(d.dstd, d.dstd + (size_t)ntypes * nnei * 4);
  compressed = d.tables != nullptr;
  tab_off.assign(nnets, 0);
  nspline.assign(nnets, 0);
  if (compressed) {
    if (!d.nspline || !d.table_info) throw std::invalid_argument("tables need nspline and table_info");
    nspline.assign(d.nspline, d.nspline + nnets);
    long long tot = 0;
    for (int n = 0; n < nnets; ++n) {
      tab_off[n] = tot;
      tot += (long long)d.nspline[n] * 6 * M;
    }
    table_host.resize(tot);
    for (int n = 0; n < nnets; ++n)
      std::memcpy(table_host.data() + tab_off[n], d.tables[n], sizeof(double) * (size_t)d.nspline[n] * 6 * M);
    info_host.assign(d.table_info, d.table_info + nnets * 6);
    check_table_rows(info_host.data(), d.nspline, nnets);
  } else {
    // uncompressed: the embedding nets themselves (FP64 MFMA operands)
    enl = d.embed_nlayers;
    if (enl <= 0 || !d.embed_width || !d.embed_w || !d.embed_b || !d.embed_act || !d.embed_resnet)
      throw std::invalid_argument("a model without tables needs its embedding nets");
    ewidth.assign(d.embed_width, d.embed_width + enl + 1);
    if (ewidth[0] != 1 || ewidth[enl] != M) throw std::invalid_argument("embedding widths must run from 1 to M");
    for (int w : ewidth)
      if (w <= 0) throw std::invalid_argument("embedding widths must be positive");
    eact.assign(d.embed_act, d.embed_act + enl);
    eres.assign(enl, 0);
    for (int l = 0; l < enl; ++l)
      if (d.embed_resnet[l]) {
        if (ewidth[l] == ewidth[l + 1])
          eres[l] = 1;
        else if (2 * ewidth[l] == ewidth[l + 1])
          eres[l] = 2;
      }
    const size_t ne = (size_t)nnets * enl;
    ew.assign(ne, nullptr);
    ewt.assign(ne, nullptr);
    eb.assign(ne, nullptr);
    eidt.assign(ne, nullptr);
    std::vector<double*> top(nnets, nullptr);
    for (int n = 0; n < nnets; ++n)
      for (int l = 0; l < enl; ++l) {
        const size_t k = (size_t)n * enl + l;
        const int ni = ewidth[l], no = ewidth[l + 1];
        const std::vector<double> ws = fit64_swizzle(d.embed_w[k], ni, no);
        ew[k] = dupload(ws.data(), ws.size());
        std::vector<double> wt((size_t)ni * no);
        for (int i = 0; i < ni; ++i)
          for (int o = 0; o < no; ++o) wt[(size_t)o * ni + i] = d.embed_w[k][(size_t)i * no + o];
        const std::vector<double> wts = fit64_swizzle(wt.data(), no, ni);
        ewt[k] = dupload(wts.data(), wts.size());
        eb[k] = dupload(d.embed_b[k], (size_t)no);
        if (d.embed_idt && d.embed_idt[k]) eidt[k] = dupload(d.embed_idt[k], (size_t)no);
        if (l == enl - 1) top[n] = eidt[k];
      }
    d_eidt_top = dupload(top.data(), top.size());
  }
  net_of.assign(ntypes * ntypes, 0);
  for (int ti = 0; ti < ntypes; ++ti)
    for (int tj = 0; tj < ntypes; ++tj) net_of[ti * ntypes + tj] = type_one_side ? tj : tj * ntypes + ti;
  nlayers = d.fit_nlayers;
  width.assign(d.fit_width, d.fit_width + nlayers + 1);
  if (width[0] != M * A) throw std::invalid_argument("fitting input width must be M*A");
  if (width[nlayers] != 1) throw std::invalid_argument("fitting output width must be 1");
  act.assign(d.fit_act, d.fit_act + nlayers);
  res_mode.assign(nlayers, 0);
  for (int l = 0; l < nlayers; ++l) {
    if (d.fit_resnet[l]) {
      if (width[l] == width[l + 1])
        res_mode[l] = 1;
      else if (2 * width[l] == width[l + 1])
        res_mode[l] = 2;
    }
  }
  bias_atom_e.assign(d.bias_atom_e, d.bias_atom_e + ntypes);
  has_idt.assign(nlayers, 0);
  for (int l = 0; l < nlayers; ++l) has_idt[l] = d.fit_idt[l] != nullptr;

  // device copies
  if (compressed) {
    d_table = dupload(table_host.data(), table_host.size());
    d_tab_off = dupload(tab_off.data(), tab_off.size());
    void* infos = make_table_infos(info_host.data(), nnets, false);
    d_infos = dupload(static_cast<const char*>(infos), table_info_bytes(false) * nnets);
    free_table_infos(infos, false);
  }
  d_net_of = dupload(net_of.data(), net_of.size());
  d_sec = dupload(sec.data(), sec.size());
  d_davg = dupload(davg.data(), davg.size());
  d_dstd = dupload(dstd.data(), dstd.size());
  fit_w.resize((size_t)ntypes * nlayers);
  fit_wt.resize((size_t)ntypes * nlayers);
  fit_b.resize((size_t)ntypes * nlayers);
  fit_idt.resize((size_t)ntypes * nlayers, nullptr);
  out_w_host.resize(ntypes);
  out_b_host.resize(ntypes);
  for (int t = 0; t < ntypes; ++t)
    for (int l = 0; l < nlayers; ++l) {
      const size_t k = (size_t)t * nlayers + l;
      const int ni = width[l], no = width[l + 1];
      {
        // FP64 MFMA operands: W and W^T ([no][ni], the backward's) as B fragments
        const std::vector<double> ws = fit64_swizzle(d.fit_w[k], ni, no);
        fit_w[k] = dupload(ws.data(), ws.size());
        std::vector<double> wt((size_t)ni * no);
        for (int i = 0; i < ni; ++i)
          for (int o = 0; o < no; ++o) wt[(size_t)o * ni + i] = d.fit_w[k][(size_t)i * no + o];
        const std::vector<double> wts = fit64_swizzle(wt.data(), no, ni);
        fit_wt[k] = dupload(wts.data(), wts.size());
      }
      fit_b[k] = dupload(d.fit_b[k], (size_t)no);
      if (d.fit_idt[k]) fit_idt[k] = dupload(d.fit_idt[k], (size_t)no);
      if (l == nlayers - 1) {
        out_w_host[t].assign(d.fit_w[k], d.fit_w[k] + ni);
        out_b_host[t] = d.fit_b[k][0];
      }
    }
  build_mixed(d);
}

// FP32 tables and FP16 fitting layers of the mixed engine.  Hidden layer l of
// type t maps x [rows, kpad[l]] -> [rows, kpad[l+1]]; padded weight rows and
// columns, biases and timesteps are zero, so padded activations stay zero.
void SeaModel::build_mixed(const dpmi_sea_model_desc& d) {
  if (!compressed) {
    mixed_unsupported = "a model without compressed tables runs in the FP64 engine only";
    return;
  }
  if (nlayers < 2)
    mixed_unsupported = "the mixed engine needs at least one hidden fitting layer";
  for (int l = 0; l < nlayers; ++l)
    if (res_mode[l] == 2) mixed_unsupported = "the mixed engine does not support width-doubling residual layers";
  // fused kernel: at most 4 hidden layers of padded width <= 256 and no
  // residual on the descriptor layer; its reductions run in steps of 128, so
  // every width is then padded to a multiple of 128 (else of 64)
  const int H = nlayers - 1;
  fused = mixed_unsupported.empty() && H <= 4 && res_mode[0] == 0;
  for (int l = 1; l <= H && fused; ++l) fused = (width[l] + 127) / 128 * 128 <= 256;
  const int align = fused ? 128 : 64;
  kpad.resize(nlayers + 1);
  for (int l = 0; l <= nlayers; ++l) kpad[l] = (width[l] + align - 1) / align * align;
  std::vector<float> tf(table_host.begin(), table_host.end());
  d_table_f = dupload(tf.data(), tf.size());
  {
    std::vector<float> pk;
    std::vector<long long> pko;
    build_packed_tables(table_host.data(), tab_off.data(), nspline.data(), nnets, M, pk, pko);
    d_table_pk = dupload(pk.data(), pk.size());
    d_tab_off_pk = dupload(pko.data(), pko.size());
  }
  void* infos = make_table_infos(info_host.data(), nnets, true);
  d_infos_f = dupload(static_cast<const char*>(infos), table_info_bytes(true) * nnets);
  free_table_infos(infos, true);
  if (embed_mfma_supported(M)) {
    std::vector<_Float16> t16;
    std::vector<long long> off16;
    std::vector<float> psc;
    build_split_tables(table_host.data(), tab_off.data(), nspline.data(), info_host.data(), nnets, M, t16, off16,
                       psc);
    d_tab16 = dupload(t16.data(), t16.size());
    d_tab16_off = dupload(off16.data(), off16.size());
    d_tab16_pscale = dupload(psc.data(), psc.size());
    embed_mfma = true;
  }
  if (embed_runs_supported(M, A)) {
    std::vector<_Float16> rt;
    std::vector<long long> roff;
    std::vector<float> psc;
    build_run_tables(table_host.data(), tab_off.data(), nspline.data(), nnets, M, rt, roff, psc);
    d_rtab = dupload(rt.data(), rt.size());
    d_rtab_off = dupload(roff.data(), roff.size());
    d_rtab_pscale = dupload(psc.data(), psc.size());
    std::vector<_Float16> ft;
    std::vector<long long> foff;
    std::vector<float> fsc;
    build_fwd_run_tables(table_host.data(), tab_off.data(), nspline.data(), info_host.data(), nnets, M, ft, foff, fsc);
    d_ftab = dupload(ft.data(), ft.size());
    d_ftab_off = dupload(foff.data(), foff.size());
    d_ftab_scale = dupload(fsc.data(), fsc.size());
    embed_runs = true;
  }
  std::vector<double> shift(ntypes);
  for (int t = 0; t < ntypes; ++t) shift[t] = out_b_host[t] + bias_atom_e[t];
  d_shift = dupload(shift.data(), shift.size());
  if (!mixed_unsupported.empty()) return;
  f16.resize((size_t)ntypes * H);
  std::vector<std::vector<_Float16>> fw((size_t)ntypes * H), bw((size_t)ntypes * H), fwl((size_t)ntypes * H),
      bwl((size_t)ntypes * H);
  for (int t = 0; t < ntypes; ++t)
    for (int l = 0; l < H; ++l) {
      const size_t k = (size_t)t * nlayers + l;
      const int ni = width[l], no = width[l + 1], kp = kpad[l], np = kpad[l + 1];
      std::vector<_Float16> wt((size_t)np * kp, (_Float16)0.f), w((size_t)kp * np, (_Float16)0.f);
      std::vector<_Float16> wtl((size_t)np * kp, (_Float16)0.f), wl((size_t)kp * np, (_Float16)0.f);
      for (int i = 0; i < ni; ++i)
        for (int o = 0; o < no; ++o) {
          const double x = d.fit_w[k][(size_t)i * no + o];
          const _Float16 v = (_Float16)x, vl = (_Float16)(x - (double)v);
          wt[(size_t)o * kp + i] = v;
          w[(size_t)i * np + o] = v;
          wtl[(size_t)o * kp + i] = vl;
          wl[(size_t)i * np + o] = vl;
        }
      std::vector<float> b(np, 0.f), idt(np, 0.f), ow(np, 0.f);
      for (int o = 0; o < no; ++o) {
        b[o] = (float)d.fit_b[k][o];
        if (d.fit_idt[k]) idt[o] = (float)d.fit_idt[k][o];
      }
      Fit16Layer& L = f16[(size_t)t * H + l];
      L.wt = dupload(wt.data(), wt.size());
      L.w = dupload(w.data(), w.size());
      L.wt_lo = dupload(wtl.data(), wtl.size());
      L.w_lo = dupload(wl.data(), wl.size());
      L.b = dupload(b.data(), b.size());
      L.idt = d.fit_idt[k] ? dupload(idt.data(), idt.size()) : nullptr;
      if (l == H - 1) {
        for (int o = 0; o < no; ++o) ow[o] = (float)out_w_host[t][o];
        L.out_w = dupload(ow.data(), ow.size());
      }
      L.act = act[l];
      L.res_mode = res_mode[l];
      L.k_pad = kp;
      L.n_pad = np;
      for (const void* p : {(const void*)L.wt, (const void*)L.w, (const void*)L.wt_lo, (const void*)L.w_lo,
                            (const void*)L.b, (const void*)L.idt, (const void*)L.out_w})
        if (p) f16_mem.push_back(const_cast<void*>(p));
      fw[(size_t)t * H + l] = std::move(wt);
      bw[(size_t)t * H + l] = std::move(w);
      fwl[(size_t)t * H + l] = std::move(wtl);
      bwl[(size_t)t * H + l] = std::move(wl);
    }
  if (!fused) return;
  std::vector<FusedLayer> fl((size_t)ntypes * H);
  for (size_t k = 0; k < fl.size(); ++k) {
    const Fit16Layer& L = f16[k];
    std::vector<_Float16> sw((size_t)L.n_pad * L.k_pad);
    swizzle_bfrags(fw[k].data(), L.n_pad, L.k_pad, sw.data());
    _Float16* wf = dupload(sw.data(), sw.size());
    swizzle_bfrags(bw[k].data(), L.k_pad, L.n_pad, sw.data());
    _Float16* wb = dupload(sw.data(), sw.size());
    swizzle_bfrags(fwl[k].data(), L.n_pad, L.k_pad, sw.data());
    _Float16* wfl = dupload(sw.data(), sw.size());
    swizzle_bfrags(bwl[k].data(), L.k_pad, L.n_pad, sw.data());
    _Float16* wbl = dupload(sw.data(), sw.size());
    for (_Float16* q : {wf, wb, wfl, wbl}) f16_mem.push_back(q);
    FusedLayer& F = fl[k];
    F.wf = wf;
    F.wb = wb;
    F.wf_lo = wfl;
    F.wb_lo = wbl;
    F.b = L.b;
    F.idt = L.idt;
    F.out_w = L.out_w;
    F.act = L.act;
    F.res = L.res_mode == 1;
    F.K = L.k_pad;
    F.N = L.n_pad;
  }
  d_fused = dupload(fl.data(), fl.size());
}

SeaModel::~SeaModel() {
  dfree(d_table);
  dfree(d_tab_off);
  dfree(d_infos);
  dfree(d_net_of);
  dfree(d_sec);
  dfree(d_davg);
  dfree(d_dstd);
  for (auto& p : fit_w) dfree(p);
  for (auto& p : fit_wt) dfree(p);
  for (auto& p : fit_b) dfree(p);
  for (auto& p : fit_idt) dfree(p);
  dfree(d_table_f);
  dfree(d_table_pk);
  dfree(d_tab_off_pk);
  dfree(d_infos_f);
  dfree(d_shift);
  dfree(d_fused);
  dfree(d_tab16);
  dfree(d_tab16_off);
  dfree(d_tab16_pscale);
  dfree(d_rtab);
  dfree(d_rtab_off);
  dfree(d_rtab_pscale);
  dfree(d_ftab);
  dfree(d_ftab_off);
  dfree(d_ftab_scale);
  for (auto& p : f16_mem) dfree(p);
  for (auto* v : {&ew, &ewt, &eb, &eidt})
    for (auto& p : *v) dfree(p);
  dfree(d_eidt_top);
}

// ---------------------------------------------------------------- engine --
SeaEngine::SeaEngine(const SeaModel* m, int nframes_, const int* natoms_h, const int* atype_h, int precision_,
                     double nall_per_atom_)
    : model(m), nframes(nframes_), precision(precision_), nall_per_atom(nall_per_atom_) {
  if (precision != DPMI_FP64 && precision != DPMI_MIXED) throw std::invalid_argument("unknown precision");
  if (precision == DPMI_MIXED && !m->mixed_unsupported.empty()) throw Unsupported(m->mixed_unsupported);
  natoms.assign(natoms_h, natoms_h + nframes);
  loc_off.assign(nframes + 1, 0);
  for (int f = 0; f < nframes; ++f) loc_off[f + 1] = loc_off[f] + natoms[f];
  ntot = loc_off[nframes];
  atype.assign(atype_h, atype_h + ntot);
  // per-row frame, centre, type permutation
  std::vector<int> frame_of(ntot), center(ntot);
  for (int f = 0; f < nframes; ++f)
    for (int i = 0; i < natoms[f]; ++i) {
      frame_of[loc_off[f] + i] = f;
      center[loc_off[f] + i] = i;
    }
  const int nt = m->ntypes;
  tstart.assign(nt + 2, 0);
  std::vector<int> cnt(nt + 1, 0);
  for (int g = 0; g < ntot; ++g) {
    const int t = atype[g];
    if (t >= nt) throw std::invalid_argument("atom type >= ntypes");
    cnt[t < 0 ? nt : t]++;
  }
  for (int t = 0; t <= nt; ++t) tstart[t + 1] = tstart[t] + cnt[t];
  std::vector<int> fill(tstart.begin(), tstart.end() - 1), pos(ntot), perm(ntot);
  for (int g = 0; g < ntot; ++g) {
    const int t = atype[g] < 0 ? nt : atype[g];
    pos[g] = fill[t]++;
    perm[pos[g]] = g;
  }
  d_frame_of = dupload(frame_of.data(), ntot);
  d_center = dupload(center.data(), ntot);
  d_type = dupload(atype.data(), ntot);
  d_pos = dupload(pos.data(), ntot);
  d_perm = dupload(perm.data(), ntot);
  geom_h.resize(nframes);
  ext_need_h.assign(nframes, 0);
  DPMI_CHECK(hipHostMalloc((void**)&geom_pinned, sizeof(FrameGeom) * nframes, hipHostMallocDefault));
  DPMI_CHECK(hipHostMalloc((void**)&stat_pinned, sizeof(int) * (nframes + 1), hipHostMallocDefault));
  d_geom = dalloc<FrameGeom>(nframes);
  d_flag = dalloc<int>(1);
  DPMI_CHECK(hipMemset(d_flag, 0, sizeof(int)));  // status bits stay set until status() reads them
  d_nall = dalloc<int>(nframes);
  d_frame_ext_off = dalloc<int>(nframes);
  const size_t nn = (size_t)ntot * m->nnei;
  d_nlist = dalloc<int>(nn);
  d_ncand = dalloc<int>(ntot);
  d_cell_of = dalloc<int>(ntot);
  d_sorted_loc = dalloc<int>(ntot);
  d_cell_slot = dalloc<int>(ntot);
  d_cell_list = dalloc<int>(ntot);
  const int D = m->M * m->A;
  if (precision == DPMI_FP64) {
    d_em = dalloc<double>(nn * 4);
    // the force kernel rebuilds em_deriv and rij from the coordinates when
    // it can (the environment kernel then writes em only)
    recompute_slots = force_recompute_supported(m->nnei);
    if (!recompute_slots) {
      d_em_deriv = dalloc<double>(nn * 12);
      d_rij = dalloc<double>(nn * 3);
    }
    d_net_deriv = dalloc<double>(nn * 4);
    d_xyz = dalloc<double>((size_t)ntot * 4 * m->M);
    d_D = dalloc<double>((size_t)ntot * D);
    d_gD = dalloc<double>((size_t)ntot * D);
    DPMI_CHECK(hipMemset(d_gD, 0, sizeof(double) * (size_t)ntot * D));
    for (int l = 0; l < m->nlayers; ++l) {
      d_fy.push_back(dalloc<double>((size_t)ntot * m->width[l + 1]));
      d_fx.push_back(dalloc<double>((size_t)ntot * m->width[l + 1]));
    }
    int wmax = 0;
    for (int w : m->width) wmax = std::max(wmax, w);
    d_g0 = dalloc<double>((size_t)ntot * wmax);
    d_g1 = dalloc<double>((size_t)ntot * wmax);
    d_gz[0] = dalloc<double>((size_t)ntot * wmax);
    d_gz[1] = dalloc<double>((size_t)ntot * wmax);
    if (!m->compressed) net_setup();
  } else {
    d_em_f = dalloc<float>(nn * 4);
    // as above (76 -> 16 bytes per slot written by the environment kernel)
    recompute_slots = force_recompute_supported(m->nnei);
    if (!recompute_slots) {
      d_em_deriv_f = dalloc<float>(nn * 12);
      d_rij_f = dalloc<float>(nn * 3);
    }
    d_net_deriv_f = dalloc<float>(nn * 4);
    d_xyz_f = dalloc<float>((size_t)ntot * 4 * m->M);
    const int k0 = m->kpad[0];
    // padded descriptor columns are never written and must stay zero
    d_Dh = dalloc<_Float16>((size_t)ntot * k0);
    DPMI_CHECK(hipMemset(d_Dh, 0, sizeof(_Float16) * (size_t)ntot * k0));
    d_gD_f = dalloc<float>((size_t)ntot * k0);
    if (m->fused && m->nlayers - 1 >= 2) {
      const size_t nx1 = (size_t)ntot * m->kpad[1];
      d_fx1 = dalloc<_Float16>(nx1);
      d_fdv1 = dalloc<_Float16>(nx1);
      d_fgz0 = dalloc<_Float16>(nx1);
    }
    d_eperm = dalloc<float>(ntot);
    int kmax = 0;
    for (int l = 0; l + 1 < m->nlayers; ++l) {
      const int np = m->kpad[l + 1];
      kmax = std::max(kmax, np);
      d_y16.push_back(dalloc<float>((size_t)ntot * np));
      d_x16.push_back(l + 2 < m->nlayers ? dalloc<_Float16>((size_t)ntot * np) : nullptr);
    }
    d_ga = dalloc<float>((size_t)ntot * kmax);
    d_gb = dalloc<float>((size_t)ntot * kmax);
    const auto chip = chip_layout();
    nxcd = chip.first;
    const std::vector<int4> tiles = fitting_tile_order(tstart, nt, 64, chip.first, chip.second);
    ntiles = (int)tiles.size();
    d_tiles = dupload(tiles.data(), tiles.size());
    const std::vector<int4> tiles32 = fitting_tile_order(tstart, nt, 32, chip.first, chip.second);
    ntiles32 = (int)tiles32.size();
    d_tiles32 = dupload(tiles32.data(), tiles32.size());
    // 64-row tiles unless they would leave compute units without a tile;
    // when even 32-row tiles leave compute units idle, layer 0's products
    // spread over more workgroups in their own kernels
    fit_rows_default = ntiles < chip.second ? 32 : 64;
    if (ntiles32 < chip.second && d_fx1) fit_rows_default = 32 | kFitSplit;
    // embedding tiles: groups of kEmbedTileAtoms descriptor positions of one
    // type (atoms of negative type last, as type -1)
    std::vector<int4> et;
    for (int t = 0; t <= nt; ++t)
      for (int p = tstart[t]; p < tstart[t + 1]; p += kEmbedTileAtoms)
        et.push_back(make_int4(t < nt ? t : -1, p, std::min(kEmbedTileAtoms, tstart[t + 1] - p), 0));
    netiles = (int)et.size();
    if (netiles) d_etiles = dupload(et.data(), et.size());
  }
  // with recomputed slot derivatives the neighbour-list kernel also stores
  // the environment matrix: one stage "nlist_env"
  stage_names = std::string(recompute_slots ? "geometry;nlist_env;" : "geometry;nlist;env_mat;") +
                (precision == DPMI_MIXED && m->fused ? "embed_fwd;fitting;embed_bwd;force_virial;reduce"
                                                     : "embed_fwd;fitting_fwd;fitting_bwd;embed_bwd;force_virial;reduce");
  d_atom_e = dalloc<double>(ntot);
  d_vpart = dalloc<double>((size_t)ntot * 9);
  {
    std::vector<int> cxy, cst;
    frame_chunks(natoms.data(), nframes, cxy, cst);
    nchunks = (int)cxy.size() / 4;
    d_chunks = dupload(cxy.data(), cxy.size());
    d_chunk_start = dupload(cst.data(), cst.size());
    d_part = dalloc<double>((size_t)std::max(nchunks, 1) * 10);
    // every frame holds a chunk: its last chunk combines it (one launch)
    bool every = true;
    for (int f = 0; f < nframes; ++f) every = every && cst[f + 1] > cst[f];
    if (every) {
      d_ticket = dalloc<int>(nframes);
      DPMI_CHECK(hipMemset(d_ticket, 0, sizeof(int) * nframes));
    }
  }
  d_outw.resize(nt, nullptr);
  for (int t = 0; t < nt; ++t) d_outw[t] = dupload(m->out_w_host[t].data(), m->out_w_host[t].size());
  // keys of the accepted candidates of one centre (rows of nnei + a quarter;
  // an overflow is flagged and the step retried with twice the capacity)
  nlist_cap = nlist_row_capacity(std::max(128, m->nnei + m->nnei / 4));
  load_cached_tune();
}

SeaEngine::~SeaEngine() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (geom_pinned) (void)hipHostFree(geom_pinned);
  if (stat_pinned) (void)hipHostFree(stat_pinned);
  for (auto* p : {&d_frame_of, &d_center, &d_type, &d_pos, &d_perm, &d_flag, &d_nall, &d_frame_ext_off, &d_nlist,
                  &d_ncand, &d_cell_of, &d_sorted_loc, &d_cell_frame, &d_xcell_frame, &d_cell_cnt, &d_cell_start,
                  &d_cell_slot, &d_cell_list, &d_xcnt, &d_xstart, &d_xgcnt, &d_xgstart, &d_cell_sub, &d_type_ext,
                  &d_map_ext, &d_xframe_of_ext, &d_chunks, &d_chunk_start, &d_ticket})
    dfree(*p);
  dfree(d_cell_atoms);
  for (auto* p : {&d_em, &d_em_deriv, &d_rij, &d_net_deriv, &d_xyz, &d_D, &d_gD, &d_atom_e, &d_vpart,
                  &d_coord_ext, &d_force_ext, &d_g0, &d_g1, &d_gz[0], &d_gz[1], &d_part})
    dfree(*p);
  for (auto* p : {&d_em_f, &d_em_deriv_f, &d_rij_f, &d_net_deriv_f, &d_xyz_f, &d_gD_f, &d_eperm, &d_ga, &d_gb})
    dfree(*p);
  dfree(d_Dh);
  for (auto* p : {&d_fx1, &d_fdv1, &d_fgz0}) dfree(*p);
  dfree(d_tiles);
  dfree(d_tiles32);
  dfree(d_etiles);
  for (auto& p : d_y16) dfree(p);
  for (auto& p : d_x16) dfree(p);
  dfree(d_geom);
  for (auto& p : d_fy) dfree(p);
  for (auto& p : d_fx) dfree(p);
  for (auto& p : d_outw) dfree(p);
  for (auto* p : {&d_nx0, &d_nds, &d_ngg, &d_ndz[0], &d_ndz[1], &d_ngx[0], &d_ngx[1]}) dfree(*p);
  for (auto& p : d_ny) dfree(p);
  for (auto& p : d_nxl) dfree(p);
  dfree(d_rowbase);
  for (auto& e : events) (void)hipEventDestroy(e);
}

// Box-dependent layout: cell grids and extended capacities per frame.
void SeaEngine::prepare_geometry(const double* box_h) {
  int ncells = 0, nxcells = 0;
  long long next = 0;
  bool grid_changed = (int)cell_frame_h.size() == 0;
  for (int f = 0; f < nframes; ++f) {
    FrameGeom& g = geom_h[f];
    const FrameGeom old = g;
    geom_from_box(g, box_h + 9 * f, model->rcut);
    for (int d = 0; d < 3; ++d) grid_changed = grid_changed || g.nc[d] != old.nc[d] || g.ng[d] != old.ng[d];
    g.nloc = natoms[f];
    g.loc_off = loc_off[f];
    g.cell_off = ncells;
    g.xcell_off = nxcells;
    ncells += geom_local_cells(g);
    nxcells += geom_ext_cells(g) + 1;
    const int cap =
        nall_per_atom > 0 ? (int)std::ceil(natoms[f] * nall_per_atom) + 64 : geom_nall_estimate(g, natoms[f], 1.25);
    g.ext_cap = std::max(std::max(cap, natoms[f]), ext_need_h[f]);
    if (!grid_changed && f < (int)ext_cap_h.size()) g.ext_cap = std::max(g.ext_cap, ext_cap_h[f]);
    g.ext_off = (int)next;
    next += g.ext_cap;
  }
  if (grid_changed || next > next_tot || ncells > ncells_tot || nxcells > nx_tot) {
    DPMI_CHECK(hipDeviceSynchronize());
    drop_graph();
    std::vector<int> cf(ncells), xf(nxcells), ef(next), eoff(nframes);
    ext_cap_h.resize(nframes);
    for (int f = 0; f < nframes; ++f) {
      const FrameGeom& g = geom_h[f];
      std::fill(cf.begin() + g.cell_off, cf.begin() + g.cell_off + geom_local_cells(g), f);
      std::fill(xf.begin() + g.xcell_off, xf.begin() + g.xcell_off + geom_ext_cells(g) + 1, f);
      std::fill(ef.begin() + g.ext_off, ef.begin() + g.ext_off + g.ext_cap, f);
      eoff[f] = g.ext_off;
      ext_cap_h[f] = g.ext_cap;
    }
    for (auto* p : {&d_cell_frame, &d_xcell_frame, &d_cell_cnt, &d_cell_start, &d_xcnt, &d_xstart,
                    &d_xgcnt, &d_xgstart, &d_cell_sub, &d_type_ext, &d_map_ext, &d_xframe_of_ext})
      dfree(*p);
    dfree(d_cell_atoms);
    dfree(d_coord_ext);
    dfree(d_force_ext);
    d_cell_frame = dupload(cf.data(), cf.size());
    d_xcell_frame = dupload(xf.data(), xf.size());
    d_xframe_of_ext = dupload(ef.data(), ef.size());
    d_cell_cnt = dalloc<int>(ncells);
    counts_clean = false;
    d_cell_start = dalloc<int>(ncells);
    d_xcnt = dalloc<int>(nxcells);
    d_xstart = dalloc<int>(nxcells);
    d_xgcnt = dalloc<int>(nxcells);
    d_xgstart = dalloc<int>(nxcells);
    d_cell_sub = dalloc<int>((size_t)nxcells * kAxisBins);
    d_type_ext = dalloc<int>(next);
    d_map_ext = dalloc<int>(next);
    d_cell_atoms = dalloc<CellAtom>(next);
    d_coord_ext = dalloc<double>((size_t)next * 3);
    // the extended force buffer and the ghost fold serve only the row
    // kernel; the aggregated kernel adds onto the owners directly
    if (!recompute_slots) d_force_ext = dalloc<double>((size_t)next * 3);
    DPMI_CHECK(hipMemcpy(d_frame_ext_off, eoff.data(), sizeof(int) * nframes, hipMemcpyHostToDevice));
    cell_frame_h = cf;
    ncells_tot = ncells;
    nx_tot = nxcells;
    next_tot = next;
  }
}

EnvArgs SeaEngine::env_args() const {
  const SeaModel& m = *model;
  EnvArgs ea;
  ea.coord = d_coord_ext;
  ea.type = d_type_ext;
  ea.nlist = d_nlist;
  ea.row_frame = d_frame_of;
  ea.row_ext_off = d_frame_ext_off;
  ea.row_center = d_center;
  ea.nrows = ntot;
  ea.nnei = m.nnei;
  ea.ntypes = m.ntypes;
  ea.avg = m.d_davg;
  ea.std_ = m.d_dstd;
  ea.rmin = (double)m.rcut_smth;
  ea.rmax = (double)m.rcut;
  ea.tpb = tune.env_tpb;
  return ea;
}

EmbedArgs SeaEngine::embed_args() const {
  const SeaModel& m = *model;
  EmbedArgs ta;
  ta.pos = d_pos;
  ta.tab_off = m.d_tab_off;
  ta.net_of = m.d_net_of;
  ta.row_type = d_type;
  ta.sec = m.d_sec;
  ta.ntypes = m.ntypes;
  ta.nrows = ntot;
  ta.nnei = m.nnei;
  ta.M = m.M;
  ta.A = m.A;
  ta.inv_nnei = 1.0 / m.nnei;
  ta.waves = tune.fwd_waves;
  ta.waves_bwd = tune.bwd_waves;
  ta.perm = d_perm;
  ta.tab16 = m.d_tab16;
  ta.tab16_off = m.d_tab16_off;
  ta.tab16_pscale = m.d_tab16_pscale;
  ta.rtab = m.d_rtab;
  ta.rtab_off = m.d_rtab_off;
  ta.rtab_pscale = m.d_rtab_pscale;
  ta.ftab = m.d_ftab;
  ta.ftab_off = m.d_ftab_off;
  ta.ftab_scale = m.d_ftab_scale;
  ta.etiles = d_etiles;
  ta.netiles = netiles;
  ta.flag = d_flag;
  return ta;
}

// FP64 stages env_mat .. embed_bwd: fitting layers on FP64 MFMA (fit64.hip).
void SeaEngine::compute_fp64_descriptor_fitting(hipStream_t st, int& stage) {
  const SeaModel& m = *model;
  if (!recompute_slots) {  // else the neighbour-list kernel stored em
    EnvArgs ea = env_args();
    ea.em = d_em;
    ea.em_deriv = d_em_deriv;
    ea.rij = d_rij;
    launch_env_mat_a(ea, st);
    mark(st, stage++);
  }

  EmbedArgs ta = embed_args();
  ta.xyz = d_xyz;
  ta.D = d_D;
  ta.gD = d_gD;
  ta.net_deriv = d_net_deriv;
  ta.tables = m.d_table;
  ta.infos = m.d_infos;
  ta.em = d_em;
  if (m.compressed)
    launch_embed_fwd<double, double>(ta, st);
  else
    net_descriptor(st);
  mark(st, stage++);

  // fitting forward: one launch per layer over every type (FP64 MFMA)
  launch_fill_double(d_atom_e, 0.0, ntot, st);
  const int L = m.nlayers;
  for (int tg = 0; tg < m.ntypes; tg += kFit64MaxTypes) {
    const int ng = std::min(kFit64MaxTypes, m.ntypes - tg);
    Fit64Layer f;
    f.ntypes = ng;
    for (int i = 0; i < ng; ++i) {
      f.p0[i] = tstart[tg + i];
      f.rows[i] = tstart[tg + i + 1] - tstart[tg + i];
    }
    for (int l = 0; l + 1 < L; ++l) {
      const bool top = l + 2 == L;
      for (int i = 0; i < ng; ++i) {
        const size_t k = (size_t)(tg + i) * L + l;
        f.w[i] = m.fit_w[k];
        f.bias[i] = m.fit_b[k];
        f.idt[i] = m.fit_idt[k];
        f.wout[i] = top ? d_outw[tg + i] : nullptr;
      }
      f.K = m.width[l];
      f.N = m.width[l + 1];
      f.act = m.act[l];
      f.res_mode = m.res_mode[l];
      f.a = l == 0 ? d_D : d_fx[l - 1];
      f.lda = m.width[l];
      f.out0 = top ? nullptr : d_fy[l];  // the top layer's dE/dz is formed here
      f.out1 = d_fx[l];
      f.out2 = top ? d_gz[0] : nullptr;
      f.ldo = m.width[l + 1];
      launch_fit64(f, false, st);
    }
    Fit64Out o;
    o.ntypes = ng;
    for (int i = 0; i < ng; ++i) {
      o.p0[i] = f.p0[i];
      o.rows[i] = f.rows[i];
      o.w[i] = d_outw[tg + i];
      o.b_out[i] = m.out_b_host[tg + i];
      o.bias[i] = m.bias_atom_e[tg + i];
    }
    o.x = L >= 2 ? d_fx[L - 2] : d_D;
    o.K = m.width[L - 1];
    o.perm = d_perm;
    o.atom_e = d_atom_e;
    launch_fit64_out(o, st);
  }
  mark(st, stage++);

  // fitting backward: dE/dD.  Each layer's product reads dE/dz of the layer
  // (formed by the previous launch) and writes dE/dz of the layer below, plus
  // the raw dE/dx the layer below needs for its residual.
  for (int tg = 0; tg < m.ntypes; tg += kFit64MaxTypes) {
    const int ng = std::min(kFit64MaxTypes, m.ntypes - tg);
    Fit64Layer f;
    f.ntypes = ng;
    for (int i = 0; i < ng; ++i) {
      f.p0[i] = tstart[tg + i];
      f.rows[i] = tstart[tg + i + 1] - tstart[tg + i];
    }
    if (L == 1) {  // no hidden layer: dE/dD = w_out on every row
      for (int i = 0; i < ng; ++i)
        if (f.rows[i] > 0) launch_fill_rows(d_gD + f.p0[i] * m.width[0], d_outw[tg + i], f.rows[i], m.width[0], st);
      continue;
    }
    int zb = 0, gb = 0;
    const double* graw = nullptr;  // dE/dx of this layer's output (the top one: w_out)
    for (int l = L - 2; l >= 0; --l) {
      const bool top = l + 2 == L;
      for (int i = 0; i < ng; ++i) {
        const size_t k = (size_t)(tg + i) * L + l;
        f.w[i] = m.fit_wt[k];
        f.wout[i] = top ? d_outw[tg + i] : nullptr;
        f.idt_next[i] = l > 0 ? m.fit_idt[k - 1] : nullptr;
      }
      f.K = m.width[l + 1];
      f.N = m.width[l];
      f.res_mode = m.res_mode[l];
      f.a = d_gz[zb];
      f.lda = m.width[l + 1];
      f.g = graw;
      f.ldg = m.width[l + 1];
      f.ldo = m.width[l];
      if (l == 0) {
        f.out0 = d_gD;
        f.out1 = nullptr;
      } else {
        f.out0 = m.res_mode[l - 1] ? (gb ? d_g1 : d_g0) : nullptr;
        f.out1 = d_gz[zb ^ 1];
        f.act_next = m.act[l - 1];
        f.y = d_fy[l - 1];
        f.ldy = m.width[l];
      }
      launch_fit64(f, true, st);
      graw = f.out0;
      gb ^= 1;
      zb ^= 1;
    }
  }
  mark(st, stage++);

  if (m.compressed)
    launch_embed_bwd<double, double>(ta, st);
  else
    net_adjoint(st);
  mark(st, stage++);

}

// ------------------------------------------------- uncompressed embedding --
long long SeaEngine::net_chunk_bytes = 16LL << 30;

// Chunks of descriptor positions whose neighbour slots fit the row buffers
// (net_chunk_bytes), and the row of every (atom, neighbour block) inside its
// chunk: rows run net by net, the atoms of a net in position order.
void SeaEngine::net_setup() {
  const SeaModel& m = *model;
  const int nt = m.ntypes, L = m.enl;
  int wmax = 0, wsum = 0;
  for (int l = 1; l <= L; ++l) {
    wmax = std::max(wmax, m.ewidth[l]);
    wsum += m.ewidth[l];
  }
  // x0, y and x per layer, dE/dG, two dE/dz and two dE/dx buffers, dE/dx0
  const long long per_row = 8LL * (2 + 2 * wsum + m.M + 4 * wmax);
  const long long cap = std::max<long long>(m.nnei, net_chunk_bytes / per_row);
  std::vector<int> perm_h(ntot);
  {
    std::vector<int> fill(tstart.begin(), tstart.end() - 1);
    for (int g = 0; g < ntot; ++g) perm_h[fill[atype[g] < 0 ? nt : atype[g]]++] = g;
  }
  std::vector<int> rowbase((size_t)ntot * nt, -1);
  net_chunks.clear();
  net_rows_max = 0;
  int q = 0;
  while (q < ntot) {
    NetChunk c;
    c.q0 = q;
    long long rows = 0;
    while (q < ntot && (rows + m.nnei <= cap || rows == 0)) {
      if (atype[perm_h[q]] >= 0) rows += m.nnei;
      ++q;
    }
    c.q1 = q;
    c.p0.assign(m.nnets, 0);
    c.rows.assign(m.nnets, 0);
    long long off = 0;
    for (int n = 0; n < m.nnets; ++n) {
      const int tj = m.type_one_side ? n : n / nt, ti = m.type_one_side ? -1 : n % nt;
      c.p0[n] = off;
      for (int p = c.q0; p < c.q1; ++p) {
        const int g = perm_h[p], t = atype[g];
        if (t < 0 || (ti >= 0 && t != ti)) continue;
        rowbase[(size_t)g * nt + tj] = (int)off;
        off += m.sel[tj];
      }
      c.rows[n] = off - c.p0[n];
    }
    if (off >= (1LL << 31)) throw std::invalid_argument("uncompressed embedding: chunk rows exceed 2^31");
    net_rows_max = std::max(net_rows_max, off);
    net_chunks.push_back(std::move(c));
  }
  d_rowbase = dupload(rowbase.data(), rowbase.size());
  const size_t R = (size_t)std::max<long long>(net_rows_max, 1);
  d_nx0 = dalloc<double>(R);
  d_nds = dalloc<double>(R);
  d_ngg = dalloc<double>(R * m.M);
  for (int k = 0; k < 2; ++k) {
    d_ndz[k] = dalloc<double>(R * wmax);
    d_ngx[k] = dalloc<double>(R * wmax);
  }
  for (int l = 0; l < L; ++l) {
    d_ny.push_back(dalloc<double>(R * m.ewidth[l + 1]));
    d_nxl.push_back(dalloc<double>(R * m.ewidth[l + 1]));
  }
}

// The embedding net on the chunk's rows: x0 -> ... -> G = d_nxl[L - 1].
void SeaEngine::net_forward(const NetChunk& c, hipStream_t st) {
  const SeaModel& m = *model;
  const int L = m.enl;
  NetArgs na;
  na.em = d_em;
  na.perm = d_perm;
  na.type = d_type;
  na.rowbase = d_rowbase;
  na.sec = m.d_sec;
  na.ntypes = m.ntypes;
  na.nnei = m.nnei;
  launch_net_gather(na, c.q0, c.q1, d_nx0, st);
  for (int n0 = 0; n0 < m.nnets; n0 += kFit64MaxTypes) {
    const int ng = std::min(kFit64MaxTypes, m.nnets - n0);
    Fit64Layer f;
    f.ntypes = ng;
    for (int i = 0; i < ng; ++i) {
      f.p0[i] = c.p0[n0 + i];
      f.rows[i] = c.rows[n0 + i];
    }
    for (int l = 0; l < L; ++l) {
      for (int i = 0; i < ng; ++i) {
        const size_t k = (size_t)(n0 + i) * L + l;
        f.w[i] = m.ew[k];
        f.bias[i] = m.eb[k];
        f.idt[i] = m.eidt[k];
      }
      f.K = m.ewidth[l];
      f.N = m.ewidth[l + 1];
      f.act = m.eact[l];
      f.res_mode = m.eres[l];
      f.a = l == 0 ? d_nx0 : d_nxl[l - 1];
      f.lda = m.ewidth[l];
      f.out0 = d_ny[l];
      f.out1 = d_nxl[l];
      f.ldo = m.ewidth[l + 1];
      launch_fit64(f, false, st);
    }
  }
}

static NetArgs net_args_of(const SeaModel& m) {
  NetArgs na;
  na.sec = m.d_sec;
  na.net_of = m.d_net_of;
  na.ntypes = m.ntypes;
  na.nnei = m.nnei;
  na.M = m.M;
  na.A = m.A;
  na.inv_nnei = 1.0 / m.nnei;
  return na;
}

// xyz and D of every atom (se_a.py:826-851 with the net evaluated for every
// slot, padding included).
void SeaEngine::net_descriptor(hipStream_t st) {
  NetArgs na = net_args_of(*model);
  na.em = d_em;
  na.perm = d_perm;
  na.pos = d_pos;
  na.type = d_type;
  na.rowbase = d_rowbase;
  na.xyz = d_xyz;
  na.D = d_D;
  for (const NetChunk& c : net_chunks) {
    net_forward(c, st);
    launch_net_contract(na, d_nxl[model->enl - 1], c.q0, c.q1, st);
  }
}

// net_deriv from dE/dD: dE/dG and the em part per slot, then dE/dG back
// through the net to its input em[g, j, 0] (layer products on W^T).  With
// more than one chunk the chunk's forward activations are recomputed.
void SeaEngine::net_adjoint(hipStream_t st) {
  const SeaModel& m = *model;
  const int L = m.enl;
  NetArgs na = net_args_of(m);
  na.em = d_em;
  na.perm = d_perm;
  na.pos = d_pos;
  na.type = d_type;
  na.rowbase = d_rowbase;
  na.xyz = d_xyz;
  na.gD = d_gD;
  na.net_deriv = d_net_deriv;
  for (const NetChunk& c : net_chunks) {
    if (net_chunks.size() > 1) net_forward(c, st);
    launch_net_adjoint(na, d_nxl[L - 1], d_ny[L - 1], m.d_eidt_top, m.eact[L - 1], c.q0, c.q1, d_ngg, d_ndz[0],
                       st);
    for (int n0 = 0; n0 < m.nnets; n0 += kFit64MaxTypes) {
      const int ng = std::min(kFit64MaxTypes, m.nnets - n0);
      Fit64Layer f;
      f.ntypes = ng;
      for (int i = 0; i < ng; ++i) {
        f.p0[i] = c.p0[n0 + i];
        f.rows[i] = c.rows[n0 + i];
      }
      int zb = 0, gb = 0;
      const double* graw = d_ngg;  // dE/d(output) of the layer
      for (int l = L - 1; l >= 0; --l) {
        for (int i = 0; i < ng; ++i) {
          const size_t k = (size_t)(n0 + i) * L + l;
          f.w[i] = m.ewt[k];
          f.idt_next[i] = l > 0 ? m.eidt[k - 1] : nullptr;
        }
        f.K = m.ewidth[l + 1];
        f.N = m.ewidth[l];
        f.res_mode = m.eres[l];
        f.a = d_ndz[zb];
        f.lda = m.ewidth[l + 1];
        f.g = graw;
        f.ldg = m.ewidth[l + 1];
        f.ldo = m.ewidth[l];
        if (l == 0) {
          f.out0 = d_nds;
          f.out1 = nullptr;
        } else {
          f.out0 = m.eres[l - 1] ? d_ngx[gb] : nullptr;
          f.out1 = d_ndz[zb ^ 1];
          f.act_next = m.eact[l - 1];
          f.y = d_ny[l - 1];
          f.ldy = m.ewidth[l];
        }
        launch_fit64(f, true, st);
        graw = f.out0;
        gb ^= 1;
        zb ^= 1;
      }
    }
    launch_net_slope(na, d_nds, c.q0, c.q1, st);
  }
}

// Mixed stages env_mat .. embed_bwd: FP32 env matrix (FP64 geometry), FP32
// table embedding and contraction, FP16 descriptor, FP16 MFMA fitting GEMMs
// with FP32 accumulation and FP32 activations/gradients.
void SeaEngine::compute_mixed_descriptor_fitting(hipStream_t st, int& stage) {
  const SeaModel& m = *model;
  if (!recompute_slots) {  // else the neighbour-list kernel stored em
    launch_env_mat_a_f32(env_args(), d_em_f, d_em_deriv_f, d_rij_f, st);
    mark(st, stage++);
  }

  const int k0 = m.kpad[0];
  EmbedArgs ta = embed_args();
  ta.xyz = d_xyz_f;
  ta.D = d_Dh;
  ta.gD = d_gD_f;
  ta.net_deriv = d_net_deriv_f;
  ta.tables = m.d_table_f;
  ta.infos = m.d_infos_f;
  ta.em = d_em_f;
  ta.ldD = k0;
  ta.packed = m.d_table_pk;
  ta.packed_off = m.d_tab_off_pk;
  const bool mfma_fwd = m.embed_mfma && (tune.embed_mfma & 1), mfma_bwd = m.embed_mfma && (tune.embed_mfma & 2);
  const bool runs_bwd = m.embed_runs && (tune.embed_mfma & 4);
  const bool runs_fwd = m.embed_runs && (tune.embed_mfma & 8);
  auto adjoint = [&] {
    if (runs_bwd)
      launch_embed_bwd_runs(ta, st);
    else if (mfma_bwd)
      launch_embed_bwd_mfma(ta, st);
    else
      launch_embed_bwd<float, float>(ta, st);
  };
  if (runs_fwd)
    launch_embed_fwd_runs(ta, st);
  else if (mfma_fwd)
    launch_embed_fwd_mfma(ta, st);
  else
    launch_embed_fwd<float, _Float16>(ta, st);
  mark(st, stage++);

  const int H = m.nlayers - 1;
  if (m.fused) {
    FusedFitArgs fa;
    fa.D = d_Dh;
    fa.ldD = k0;
    fa.gD = d_gD_f;
    fa.ldG = k0;
    const int knob = tune.fit_rows ? tune.fit_rows : fit_rows_default;
    const int rows = knob & 255;
    const bool split = (knob & kFitSplit) && d_fx1;
    fa.tiles = rows == 32 ? d_tiles32 : d_tiles;
    fa.layers = m.d_fused;
    fa.perm = d_perm;
    fa.atom_e = d_atom_e;
    fa.shift = m.d_shift;
    fa.H = H;
    // atoms of negative type own no tile: their energy is zero
    if (tstart[m.ntypes + 1] > tstart[m.ntypes]) launch_fill_double(d_atom_e, 0.0, ntot, st);
    if (split) {
      fa.x1 = d_fx1;
      fa.dv1 = d_fdv1;
      fa.gz0 = d_fgz0;
      fa.ldX = m.kpad[1];
      fa.nxcd = nxcd;
      FusedLayer l0;
      l0.K = k0;
      l0.N = m.kpad[1];
      launch_fit_split(fa, rows == 32 ? ntiles32 : ntiles, rows, l0, st);
    } else {
      launch_fit_fused(fa, rows == 32 ? ntiles32 : ntiles, rows, st, rows == 64 && (knob & kFitWide));
    }
    mark(st, stage++);
    adjoint();
    mark(st, stage++);
    return;
  }
  launch_fill_float(d_eperm, 0.f, ntot, st);
  for (int t = 0; t < m.ntypes; ++t) {
    const long long p0 = tstart[t], n = tstart[t + 1] - tstart[t];
    if (n == 0) continue;
    const _Float16* x = d_Dh + (size_t)p0 * k0;
    int ldx = k0;
    for (int l = 0; l < H; ++l) {
      const Fit16Layer& L = m.f16[(size_t)t * H + l];
      const int ldo = m.kpad[l + 1];
      const bool last = l == H - 1;
      _Float16* xo = last ? nullptr : d_x16[l] + (size_t)p0 * ldo;
      launch_fit16_fwd(L, x, ldx, (int)n, L.res_mode == 1 ? x : nullptr, ldx, d_y16[l] + (size_t)p0 * ldo, ldo, xo,
                       ldo, last ? d_eperm + p0 : nullptr, st);
      x = xo;
      ldx = ldo;
    }
  }
  launch_energy_out(d_eperm, d_pos, d_type, m.d_shift, ntot, d_atom_e, st);
  mark(st, stage++);

  for (int t = 0; t < m.ntypes; ++t) {
    const long long p0 = tstart[t], n = tstart[t + 1] - tstart[t];
    if (n == 0) continue;
    const float* g = nullptr;  // dE/dx of the current layer output
    int ldg = 0;
    float* bufs[2] = {d_ga, d_gb};
    int nb = 0;
    for (int l = H - 1; l >= 0; --l) {
      const Fit16Layer& L = m.f16[(size_t)t * H + l];
      const int ldo = m.kpad[l + 1], ldi = m.kpad[l];
      const bool top = l == H - 1;
      float* gin = l == 0 ? d_gD_f + (size_t)p0 * k0 : bufs[nb] + (size_t)p0 * ldi;
      launch_fit16_bwd(L, top ? L.out_w : g, ldg, top, d_y16[l] + (size_t)p0 * ldo, ldo, (int)n, L.res_mode == 1,
                       gin, ldi, st);
      g = gin;
      ldg = ldi;
      nb ^= 1;
    }
  }
  mark(st, stage++);

  adjoint();
  mark(st, stage++);
}

// Stage events are kept per call (up to kMaxTimedCalls) so that a timed loop
// of many asynchronous calls can be read back once at the end.
constexpr int kMaxTimedCalls = 4096;
constexpr int kMaxStages = 16;

void SeaEngine::mark(hipStream_t st, int stage) {
  if (!timing || timed_calls >= kMaxTimedCalls) return;
  const size_t k = (size_t)timed_calls * kMaxStages + stage;
  while (events.size() <= k) {
    hipEvent_t e;
    DPMI_CHECK(hipEventCreate(&e));
    events.push_back(e);
  }
  DPMI_CHECK(hipEventRecord(events[k], st));
}


void SeaEngine::drop_graph() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  graph_exec = nullptr;
  graph = nullptr;
}

// The reference grows its ghost and neighbour buffers and retries the step
// (prod_env_mat_multi_device.cc:2015-2028, 2069-2083).  Here an eager step
// with growth on reads back its status word and per-frame nall; a ghost
// overflow raises that frame's extended capacity to the nall it reported
// (plus a margin), a neighbour-row overflow doubles the row capacity (up to
// 4096 candidates), and the step is run again.  Recovered bits are not
// reported by status(); anything else stays set for it.
void SeaEngine::compute(const double* coord, const double* box_h, double* energy, double* force, double* virial,
                        double* atom_energy, double* atom_virial, hipStream_t st) {
  constexpr int kMaxRegrow = 8;
  for (int attempt = 0;; ++attempt) {
    enqueue_step(coord, box_h, energy, force, virial, atom_energy, atom_virial, st);
    if (!grow || capturing) return;
    DPMI_CHECK(hipMemcpyAsync(stat_pinned, d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipMemcpyAsync(stat_pinned + 1, d_nall, sizeof(int) * nframes, hipMemcpyDeviceToHost, st));
    DPMI_CHECK(hipStreamSynchronize(st));
    const int flag = stat_pinned[0];
    if (flag) DPMI_CHECK(hipMemset(d_flag, 0, sizeof(int)));
    int left = flag;
    bool again = false;
    if ((flag & kFlagGhostCapacity) && attempt < kMaxRegrow) {
      for (int f = 0; f < nframes; ++f)
        if (stat_pinned[1 + f] > geom_h[f].ext_cap)
          ext_need_h[f] = std::max(ext_need_h[f], stat_pinned[1 + f] + stat_pinned[1 + f] / 16 + 64);
      left &= ~kFlagGhostCapacity;
      again = true;
    }
    if ((flag & kFlagNlistCapacity) && attempt < kMaxRegrow && nlist_cap < 4096) {
      nlist_cap = nlist_row_capacity(2 * nlist_cap);
      left &= ~kFlagNlistCapacity;
      again = true;
    }
    sticky |= left;
    if (!again) return;
    ++regrows;
  }
}

void SeaEngine::enqueue_step(const double* coord, const double* box_h, double* energy, double* force,
                             double* virial, double* atom_energy, double* atom_virial, hipStream_t st) {
  const SeaModel& m = *model;
  prepare_geometry(box_h);
  last_stream = st;
  int stage = 0;
  mark(st, stage++);
  // upload the per-frame layout only when it changed; the pinned staging
  // buffer is rewritten only after the previous upload has completed.  A new
  // layout (box) invalidates a captured graph, which holds the old one.
  if (!geom_uploaded || std::memcmp(geom_pinned, geom_h.data(), sizeof(FrameGeom) * nframes) != 0) {
    DPMI_CHECK(hipStreamSynchronize(st));
    if (!capturing && !graph_dbox) drop_graph();  // a device-box graph re-reads its box
    std::memcpy(geom_pinned, geom_h.data(), sizeof(FrameGeom) * nframes);
    DPMI_CHECK(hipMemcpyAsync(d_geom, geom_pinned, sizeof(FrameGeom) * nframes, hipMemcpyHostToDevice, st));
    geom_uploaded = true;
  }
  // a step captured for a box in device memory re-reads it on every replay
  if (capturing && graph_box) launch_geom_refresh(graph_box, d_geom, nframes, m.rcut, d_flag, st);
  // force/virial arguments (the aggregated force kernel adds onto the owner
  // rows of `force`, which the wrap kernel clears)
  ForceRowsArgs fa;
  if (precision == DPMI_FP64) {
    fa.net_deriv = d_net_deriv;
    fa.em_deriv = d_em_deriv;  // null: recomputed
    fa.rij = d_rij;
  } else {
    fa.net_deriv = d_net_deriv_f;
    fa.em_deriv = d_em_deriv_f;  // null: recomputed
    fa.rij = d_rij_f;
    fa.fp32 = true;
  }
  fa.coord_ext = d_coord_ext;
  fa.type_ext = d_type_ext;
  fa.dstd = m.d_dstd;
  fa.rmin = (double)m.rcut_smth;
  fa.rmax = (double)m.rcut;
  fa.nlist = d_nlist;
  fa.row_frame = d_frame_of;
  fa.row_center = d_center;
  fa.frame_ext_off = d_frame_ext_off;
  fa.nrows = ntot;
  fa.nnei = m.nnei;
  fa.force_ext = d_force_ext;
  fa.vpart = d_vpart;
  fa.sorted_loc = d_sorted_loc;
  fa.geom = d_geom;
  fa.rows = tune.force_rows;
  fa.map_ext = d_map_ext;
  fa.force = force;
  fa.atom_virial = atom_virial;
  const bool owner_force = force_rows_aggregated(fa);
  if (atom_virial && !owner_force) throw Unsupported("per-atom virials need the LDS-aggregated force kernel");

  CellWork w;
  w.nframes = nframes;
  w.ntot = ntot;
  w.max_nloc = *std::max_element(natoms.begin(), natoms.end());
  w.ncells_tot = ncells_tot;
  w.nx_tot = nx_tot;
  w.do_wrap = 1;
  w.coord_in = coord;
  w.type_in = d_type;
  w.frame_of = d_frame_of;
  w.cell_frame = d_cell_frame;
  w.xcell_frame = d_xcell_frame;
  w.geom = d_geom;
  w.coord_ext = d_coord_ext;
  w.type_ext = d_type_ext;
  w.map_ext = d_map_ext;
  w.cell_atoms = d_cell_atoms;
  w.cell_of = d_cell_of;
  w.cell_cnt = d_cell_cnt;
  w.cell_start = d_cell_start;
  w.cell_slot = d_cell_slot;
  w.cell_list = d_cell_list;
  w.counts_clear = counts_clean;
  w.sorted_loc = d_sorted_loc;
  w.xcnt = d_xcnt;
  w.xstart = d_xstart;
  w.xgcnt = d_xgcnt;
  w.xgstart = d_xgstart;
  w.cell_sub = d_cell_sub;
  w.nall = d_nall;
  w.flag = d_flag;
  w.force_clear = owner_force ? force : nullptr;
  counts_clean = false;
  launch_build_extended(w, st);
  counts_clean = true;
  mark(st, stage++);

  NlistArgs na;
  na.cap = nlist_cap;
  na.sec = m.d_sec;
  na.ntypes = m.ntypes;
  na.nnei = m.nnei;
  na.rc2 = (double)(m.rcut * m.rcut);
  na.rc2f = m.rcut * m.rcut;
  na.nlist = d_nlist;
  na.ncand = d_ncand;
  if (recompute_slots) {
    // the environment matrix without derivatives comes with the rows
    na.avg = m.d_davg;
    na.stdv = m.d_dstd;
    na.rmin = (double)m.rcut_smth;
    na.rmax = (double)m.rcut;
    if (precision == DPMI_FP64)
      na.em_d = d_em;
    else
      na.em_f = d_em_f;
  }
  launch_cell_nlist(w, na, st);
  mark(st, stage++);

  if (precision == DPMI_FP64)
    compute_fp64_descriptor_fitting(st, stage);
  else
    compute_mixed_descriptor_fitting(st, stage);

  if (owner_force) {
    if (atom_virial) launch_fill_double(atom_virial, 0.0, (long long)ntot * 9, st);
    launch_force_virial_rows(fa, st);
  } else {
    launch_fill_double(d_force_ext, 0.0, next_tot * 3, st);
    launch_force_virial_rows(fa, st);
    launch_fill_double(force, 0.0, (long long)ntot * 3, st);
    launch_fold_ghost_forces(d_force_ext, d_map_ext, d_xframe_of_ext, d_geom, d_nall, next_tot, force, st);
  }
  mark(st, stage++);
  launch_frame_reduce(d_atom_e, d_vpart, d_chunks, nchunks, d_chunk_start, nframes, d_part, d_ticket, energy, virial,
                      st);
  if (atom_energy)
    launch_copy_double(atom_energy, d_atom_e, ntot, st);
  mark(st, stage++);
  nstages = stage - 1;
  if (timing && timed_calls < kMaxTimedCalls) ++timed_calls;
}

// forward_lower (make_model.py:372-453) on the lib's external-list
// conventions (prod_env_mat_multi_device.cc:2313-2396): the caller owns the
// ghosts and the candidate rows; the rows are formatted as format_nlist does
// (fmt_nlist.cc:99-177: FP32 distance test at rcut, so an rcut + skin list is
// cut, sorted by (type, r^2, index), truncated per type at sel), the
// environment rows are computed from the caller's coordinates, and every
// force and per-atom virial term stays on the extended atom it belongs to,
// as extended_force / extended_virial (transform_output.py:70-101).
void SeaEngine::compute_lower(const double* coord_ext, const int* atype_ext, int nall, int inum, const int* ilist,
                              const int* numneigh, const int* jlist, int jstride, double* energy, double* force_ext,
                              double* virial, double* atom_energy, double* atom_virial_ext, hipStream_t st) {
  const SeaModel& m = *model;
  if (nframes != 1) throw std::invalid_argument("the lower interface takes engines of one frame");
  if (nall < ntot) throw std::invalid_argument("nall must be at least the engine's local atom count");
  if (inum != ntot) throw std::invalid_argument("inum must equal the number of local atoms");
  if (jstride < 0 || jstride > 4096) throw std::invalid_argument("jstride (the longest candidate row) must be <= 4096");
  if (!recompute_slots) throw Unsupported("the lower interface needs nnei < 4096");
  last_stream = st;
  if (!geom_uploaded) {  // one frame at offset 0 (the force kernel reads the frame layout)
    std::memset(geom_pinned, 0, sizeof(FrameGeom));
    DPMI_CHECK(hipMemcpyAsync(d_geom, geom_pinned, sizeof(FrameGeom), hipMemcpyHostToDevice, st));
    DPMI_CHECK(hipMemsetAsync(d_frame_ext_off, 0, sizeof(int), st));
    geom_uploaded = true;
  }
  launch_lower_prep(force_ext, atom_virial_ext, nall, atype_ext, d_type, ntot, m.ntypes, d_flag, st);
  launch_fill_int(d_nlist, -1, (long long)ntot * m.nnei, st);
  NlistArgs na;
  na.cap = nlist_row_capacity(std::max(jstride, 64));
  na.sec = m.d_sec;
  na.ntypes = m.ntypes;
  na.nnei = m.nnei;
  na.rc2f = m.rcut * m.rcut;
  na.nlist = d_nlist;
  launch_row_format(inum, ilist, numneigh, jlist, jstride, coord_ext, atype_ext, ntot, nall, na, d_flag, st);
  EnvArgs ea;
  ea.coord = coord_ext;
  ea.type = atype_ext;
  ea.nlist = d_nlist;
  ea.strided_nloc = ntot;
  ea.nrows = ntot;
  ea.nnei = m.nnei;
  ea.nall_stride = nall;
  ea.ntypes = m.ntypes;
  ea.avg = m.d_davg;
  ea.std_ = m.d_dstd;
  ea.rmin = (double)m.rcut_smth;
  ea.rmax = (double)m.rcut;
  ea.tpb = tune.env_tpb;
  int stage = 2;
  if (precision == DPMI_FP64) {
    ea.em = d_em;
    launch_env_mat_a(ea, st);
    compute_fp64_descriptor_fitting(st, stage);
  } else {
    launch_env_mat_a_f32(ea, d_em_f, nullptr, nullptr, st);
    compute_mixed_descriptor_fitting(st, stage);
  }
  ForceRowsArgs fa;
  fa.net_deriv = precision == DPMI_FP64 ? (const void*)d_net_deriv : (const void*)d_net_deriv_f;
  fa.fp32 = precision != DPMI_FP64;
  fa.coord_ext = coord_ext;
  fa.type_ext = d_type;  // centre types: the engine's (validated) local types
  fa.dstd = m.d_dstd;
  fa.rmin = (double)m.rcut_smth;
  fa.rmax = (double)m.rcut;
  fa.nlist = d_nlist;
  fa.row_frame = d_frame_of;
  fa.row_center = d_center;
  fa.frame_ext_off = d_frame_ext_off;
  fa.nrows = ntot;
  fa.nnei = m.nnei;
  fa.vpart = d_vpart;
  fa.sorted_loc = d_center;  // rows in the caller's order
  fa.geom = d_geom;
  fa.rows = tune.force_rows;
  fa.unfolded = true;
  fa.force = force_ext;
  fa.atom_virial = atom_virial_ext;
  launch_force_virial_rows(fa, st);
  launch_frame_reduce(d_atom_e, d_vpart, d_chunks, nchunks, d_chunk_start, nframes, d_part, d_ticket, energy, virial,
                      st);
  if (atom_energy) launch_copy_double(atom_energy, d_atom_e, ntot, st);
}

// Tuned launch shapes per (precision, log2 of the packed atom count).
namespace {
std::mutex g_tune_mu;
std::map<long long, Tune>& tune_cache() {
  static auto* m = new std::map<long long, Tune>();
  return *m;
}
long long tune_key(int precision, int ntot) {
  int b = 0;
  while ((1LL << (b + 1)) <= ntot) ++b;
  return (long long)precision * 64 + b;
}
}  // namespace

void SeaEngine::load_cached_tune() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = tune_cache().find(tune_key(precision, ntot));
  if (it != tune_cache().end()) tune = it->second;
}

// Times `reps` rounds of one step per candidate (three candidate values of
// every knob, varied together since the knobs act on different stages; the
// candidates interleaved so that clock and cache drift spread over all of
// them) and keeps, per stage, the value with the fastest single step; the
// result is cached for the size bucket.
void SeaEngine::autotune(const double* coord, const double* box_h, int reps, hipStream_t st) {
  const int env_c[3] = {128, 256, 512}, fwd_c[3] = {1, 2, 4}, bwd_c[3] = {1, 2, 4}, frc_c[3] = {8, 32, 64},
            fit_c[3] = {32 | kFitSplit, 64 | kFitWide, 32};
  // candidates 0..2: per-neighbour embedding kernels with the shapes above;
  // then, in the mixed engine, the table-space embedding kernels (mask 3)
  // and the row-run forward and adjoint (mask 12), whose launch shapes are
  // fixed
  int masks[5] = {0, 0, 0, 0, 0};
  int ncand = 3;
  if (precision == DPMI_MIXED && model->embed_mfma) masks[ncand++] = 3;
  if (precision == DPMI_MIXED && model->embed_runs) masks[ncand++] = 12;
  double t_env[5], t_fwd[5], t_bwd[5], t_frc[5], t_fit[5];
  double *e = nullptr, *f = nullptr, *v = nullptr;
  DPMI_CHECK(hipMalloc(&e, sizeof(double) * nframes));
  DPMI_CHECK(hipMalloc(&f, sizeof(double) * 3 * (size_t)ntot));
  DPMI_CHECK(hipMalloc(&v, sizeof(double) * 9 * nframes));
  const bool was_timing = timing;
  std::vector<std::string> names;
  for (size_t a = 0, b; a <= stage_names.size(); a = b + 1) {
    b = stage_names.find(';', a);
    if (b == std::string::npos) b = stage_names.size();
    names.push_back(stage_names.substr(a, b - a));
  }
  auto idx = [&](const char* n) {
    for (size_t k = 0; k < names.size(); ++k)
      if (names[k] == n) return (int)k;
    return -1;
  };
  try {
    compute(coord, box_h, e, f, v, nullptr, nullptr, st);  // geometry and buffers
    std::vector<double> ms(names.size(), 0.0);
    for (int r = 0; r < std::max(reps, 1); ++r) {
      for (int c = 0; c < ncand; ++c) {
        const int k = c < 3 ? c : 1;
        tune.env_tpb = env_c[k];
        tune.fwd_waves = fwd_c[k];
        tune.bwd_waves = bwd_c[k];
        tune.force_rows = frc_c[k];
        tune.fit_rows = fit_c[k];
        tune.embed_mfma = masks[c];
        timing = true;
        reset_timing();
        compute(coord, box_h, e, f, v, nullptr, nullptr, st);
        timing_report(ms.data(), (int)ms.size(), nullptr, 0);
        auto keep = [&](double* t, const char* n) {
          const double x = idx(n) >= 0 ? ms[idx(n)] : 0.0;
          t[c] = r == 0 ? x : std::min(t[c], x);
        };
        keep(t_env, "env_mat");
        keep(t_fwd, "embed_fwd");
        keep(t_bwd, "embed_bwd");
        keep(t_frc, "force_virial");
        keep(t_fit, "fitting");
      }
    }
  } catch (...) {
    timing = was_timing;
    (void)hipFree(e);
    (void)hipFree(f);
    (void)hipFree(v);
    throw;
  }
  timing = was_timing;
  reset_timing();
  DPMI_CHECK(hipStreamSynchronize(st));
  (void)hipFree(e);
  (void)hipFree(f);
  (void)hipFree(v);
  auto best = [](const double* t) { return t[0] <= t[1] ? (t[0] <= t[2] ? 0 : 2) : (t[1] <= t[2] ? 1 : 2); };
  tune.env_tpb = env_c[best(t_env)];
  tune.fwd_waves = fwd_c[best(t_fwd)];
  tune.bwd_waves = bwd_c[best(t_bwd)];
  tune.force_rows = frc_c[best(t_frc)];
  tune.fit_rows = idx("fitting") >= 0 ? fit_c[best(t_fit)] : 0;
  tune.embed_mfma = 0;
  double fbest = t_fwd[best(t_fwd)], bbest = t_bwd[best(t_bwd)];
  int bmask = 0;
  int fmask = 0;
  for (int c = 3; c < ncand; ++c) {
    if ((masks[c] & 9) && t_fwd[c] < fbest) {
      fbest = t_fwd[c];
      fmask = masks[c] & 9;
    }
    if ((masks[c] & 6) && t_bwd[c] < bbest) {
      bbest = t_bwd[c];
      bmask = masks[c] & 6;
    }
  }
  tune.embed_mfma = fmask | bmask;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_cache()[tune_key(precision, ntot)] = tune;
}

// The captured step holds only kernel nodes (every clear is a fill kernel
// and no library is called, in either precision), so replays are exact
// repeats of an eager step on the current coordinates.
void SeaEngine::capture_device_box(const double* coord, const double* box_dev, double* energy, double* force,
                                   double* virial, double* atom_energy, hipStream_t st) {
  if (!st) throw std::invalid_argument("graph capture needs a non-default stream");
  std::vector<double> box_h((size_t)nframes * 9);
  DPMI_CHECK(hipMemcpyAsync(box_h.data(), box_dev, sizeof(double) * box_h.size(), hipMemcpyDeviceToHost, st));
  DPMI_CHECK(hipStreamSynchronize(st));
  graph_box = box_dev;
  try {
    capture(coord, box_h.data(), energy, force, virial, atom_energy, st);
  } catch (...) {
    graph_box = nullptr;
    throw;
  }
  graph_box = nullptr;
  graph_dbox = true;
}

void SeaEngine::capture(const double* coord, const double* box_h, double* energy, double* force, double* virial,
                        double* atom_energy, hipStream_t st) {
  if (!st) throw std::invalid_argument("graph capture needs a non-default stream");
  // geometry, buffers, uploads (and any capacity growth) for these boxes
  graph_dbox = false;
  compute(coord, box_h, energy, force, virial, atom_energy, nullptr, st);
  DPMI_CHECK(hipStreamSynchronize(st));
  drop_graph();
  const bool was_timing = timing;
  timing = false;
  capturing = true;
  DPMI_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    compute(coord, box_h, energy, force, virial, atom_energy, nullptr, st);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    timing = was_timing;
    capturing = false;
    throw;
  }
  capturing = false;
  timing = was_timing;
  DPMI_CHECK(hipStreamEndCapture(st, &graph));
  DPMI_CHECK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
}

void SeaEngine::launch_graph(hipStream_t st) {
  if (!graph_exec)
    throw std::invalid_argument(
        "no captured step for the current buffers and boxes (none captured, or a later step changed the box or "
        "reallocated): capture again");
  DPMI_CHECK(hipGraphLaunch(graph_exec, st));
  last_stream = st;
  if (graph_dbox) geom_uploaded = false;  // the device geometry now holds the replayed box
}

int SeaEngine::status_flags() {
  DPMI_CHECK(hipStreamSynchronize(last_stream));
  int flag = 0;
  DPMI_CHECK(hipMemcpy(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) DPMI_CHECK(hipMemset(d_flag, 0, sizeof(int)));
  flag |= sticky;
  sticky = 0;
  return flag;
}

long long SeaEngine::fetch(int frame, int what, void* dst) {
  const SeaModel& m = *model;
  DPMI_CHECK(hipStreamSynchronize(last_stream));
  if (frame < 0 || frame >= nframes) throw std::invalid_argument("frame out of range");
  int nall = 0;
  DPMI_CHECK(hipMemcpy(&nall, d_nall + frame, sizeof(int), hipMemcpyDeviceToHost));
  const FrameGeom& g = geom_h[frame];
  const size_t r0 = loc_off[frame];
  const int nloc = natoms[frame];
  const void* src = nullptr;
  long long n = 0;
  size_t es = 0;
  switch (what) {
    case 0:
      if (dst) std::memcpy(dst, &nall, sizeof(int));
      return 1;
    case 1: src = d_coord_ext + (size_t)g.ext_off * 3; n = 3LL * nall; es = 8; break;
    case 2: src = d_type_ext + g.ext_off; n = nall; es = 4; break;
    case 3: src = d_map_ext + g.ext_off; n = nall; es = 4; break;
    case 4: src = d_nlist + r0 * m.nnei; n = (long long)nloc * m.nnei; es = 4; break;
    case 5:
    case 7: {
      n = (long long)nloc * m.nnei * 4;
      if (precision == DPMI_FP64) {
        src = (what == 5 ? d_em : d_net_deriv) + r0 * m.nnei * 4;
        es = 8;
        break;
      }
      // mixed engine: FP32 storage, returned as double
      if (dst) {
        std::vector<float> tmp(n);
        DPMI_CHECK(hipMemcpy(tmp.data(), (what == 5 ? d_em_f : d_net_deriv_f) + r0 * m.nnei * 4, sizeof(float) * n,
                             hipMemcpyDeviceToHost));
        std::copy(tmp.begin(), tmp.end(), static_cast<double*>(dst));
      }
      return n;
    }
    case 6: {
      // descriptor rows are stored type-permuted (FP16 rows of stride kpad[0]
      // in the mixed engine)
      const int D = m.M * m.A;
      n = (long long)nloc * D;
      if (dst) {
        std::vector<int> pos(ntot);
        DPMI_CHECK(hipMemcpy(pos.data(), d_pos, sizeof(int) * ntot, hipMemcpyDeviceToHost));
        std::vector<_Float16> row(D);
        for (int i = 0; i < nloc; ++i) {
          double* out = static_cast<double*>(dst) + (size_t)i * D;
          if (precision == DPMI_FP64) {
            DPMI_CHECK(hipMemcpy(out, d_D + (size_t)pos[r0 + i] * D, sizeof(double) * D, hipMemcpyDeviceToHost));
          } else {
            DPMI_CHECK(hipMemcpy(row.data(), d_Dh + (size_t)pos[r0 + i] * m.kpad[0], sizeof(_Float16) * D,
                                 hipMemcpyDeviceToHost));
            for (int k = 0; k < D; ++k) out[k] = (double)row[k];
          }
        }
      }
      return n;
    }
    case 8: src = d_atom_e + r0; n = nloc; es = 8; break;
    case 9:  // extended capacity of the frame
      if (dst) std::memcpy(dst, &g.ext_cap, sizeof(int));
      return 1;
    case 10:  // capacity increases made so far
      if (dst) std::memcpy(dst, &regrows, sizeof(int));
      return 1;
    case 11:  // neighbour-row candidate capacity
      if (dst) std::memcpy(dst, &nlist_cap, sizeof(int));
      return 1;
    default: throw std::invalid_argument("unknown fetch target");
  }
  if (dst && n) DPMI_CHECK(hipMemcpy(dst, src, es * n, hipMemcpyDeviceToHost));
  return n;
}

// Average per-stage device time over the calls recorded since timing was
// (re)enabled.
int SeaEngine::timing_report(double* ms, int max_stages, char* names, int names_len) {
  if (!timing || nstages <= 0 || timed_calls == 0) return 0;
  DPMI_CHECK(hipStreamSynchronize(last_stream));
  const int n = std::min(nstages, max_stages);
  for (int s = 0; s < n; ++s) {
    double acc = 0.;
    for (int c = 0; c < timed_calls; ++c) {
      float t = 0.f;
      DPMI_CHECK(hipEventElapsedTime(&t, events[(size_t)c * kMaxStages + s], events[(size_t)c * kMaxStages + s + 1]));
      acc += t;
    }
    ms[s] = acc / timed_calls;
  }
  if (names && names_len > 0) {
    std::strncpy(names, stage_names.c_str(), names_len - 1);
    names[names_len - 1] = 0;
  }
  return n;
}

}  // namespace dpmi
