// Batched uneven-size execution: several plans (different L and K) in one hipGraph, captured once
// and replayed (cfk_multi_* in include/cfk.h).  The graph converts each group's f to FP32, then forks
// one item kernel per radial size K onto parallel branches, so the kernels of different K share the
// GPU and fill each other's tails.  Each group runs the item kernel its plan's mixed-precision tile
// selects: the VALU kernel (contract.hip) or the matrix-core one (contract_mx.hip, with or without
// the row split).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace cfk {
int launch_multi(const MultiGroupDev *groups, const int4 *items, int K, int KC, int i0, int i1, hipStream_t st);
int launch_multi_mx(const MultiGroupDev *groups, const int4 *items, int K, bool split, int i0, int i1,
                    hipStream_t st);
int launch_to_f32(const double *in, float *out, int64_t n, hipStream_t st);
int launch_pack_rows(const double *in, float *out, int N_Q, int K, int KC, int64_t B, bool natural, hipStream_t st);
bool multi_supports_k(int K);
Tile plan_tile(const cfk_plan_s *p, int precision, int64_t B = 0);
int ensure_mx_tables(cfk_plan_s *p);
}  // namespace cfk

using namespace cfk;

struct cfk_multi_s {
    std::vector<cfk_plan_s *> plans;
    int device = 0;
    hipStream_t cap = nullptr;
    hipStream_t branch[MULTI_MAX_GROUPS] = {};   // fork streams used during capture only
    hipEvent_t ev[MULTI_MAX_GROUPS + 1] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t launched = nullptr;  // recorded after each graph launch
    // the pointer/size set the current graph was captured for
    std::vector<const double *> key_f;
    std::vector<double *> key_Q;
    std::vector<int64_t> key_B;
    std::vector<int> key_acc;  // each plan's output mode (cfk_plan_set_accumulate) at capture
    std::unique_ptr<DevBuf<int4>> items{new DevBuf<int4>()};
    std::unique_ptr<DevBuf<MultiGroupDev>> groups{new DevBuf<MultiGroupDev>()};
    bool use_graph = true;
    bool prepared = false;
    // Once a graph has been captured it is neither destroyed nor captured again before the multi plan
    // is: a call with other pointers, sizes or output modes keeps the graph and the buffers it
    // references (below) and runs the same kernels by direct launches from then on.  (Destroying the
    // executable graph and capturing a new one during the plan's life crashed the host process.)
    bool frozen = false;
    std::vector<std::unique_ptr<DevBuf<int4>>> kept_items;
    std::vector<std::unique_ptr<DevBuf<MultiGroupDev>>> kept_groups;
    std::vector<std::unique_ptr<DevBuf<float>>> kept_f32;
    struct Launch {
        int K, KC, i0, i1;
        bool mx, split;  // radial stage on the matrix cores; two waves per item
    };
    std::vector<Launch> launches;  // one item kernel per K-class
    std::vector<std::unique_ptr<DevBuf<float>>> f32;
    std::vector<char> packed;  // per group: its item kernel reads packed FP32 rows
    std::vector<int> pack_kc;  // per group: the k3 chunk width of its packed rows
    int n_items = 0;
    ~cfk_multi_s() {
        if (launched) (void)hipEventSynchronize(launched);  // the last launch (graph or direct) is done
        reset();
        if (launched) (void)hipEventDestroy(launched);
        if (cap) (void)hipStreamDestroy(cap);
        for (auto &b : branch)
            if (b) (void)hipStreamDestroy(b);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    // Destroys the captured graph.  A launch of it may still be running (execute returns as soon as
    // the graph is enqueued): wait for the last launch first, since destroying an executable graph
    // in flight crashes the runtime.
    void reset() {
        if (exec && launched) (void)hipEventSynchronize(launched);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

static int multi_prepare(cfk_multi_s *m, double *const *Q, const double *const *f, const int64_t *B) {
    const int G = (int)m->plans.size();
    // the buffers below are replaced: wait for the last launch that reads them (graph or direct)
    if (m->launched) (void)hipEventSynchronize(m->launched);
    if (m->exec && !m->frozen) {
        // keep the captured graph and what it reads
        m->frozen = true;
        m->kept_items.push_back(std::move(m->items));
        m->kept_groups.push_back(std::move(m->groups));
        m->items.reset(new DevBuf<int4>());
        m->groups.reset(new DevBuf<MultiGroupDev>());
        for (auto &b : m->f32) m->kept_f32.push_back(std::move(b));
        m->f32.clear();
    }
    m->prepared = false;
    // K-classes: groups that share the item kernel -- (K, KC) on the VALU kernel, (K, split) on the
    // matrix-core one -- as each plan's mixed-precision tile selects it
    struct Cls {
        int K, KC;
        bool mx, split;
        double work = 0.0;
        std::vector<int4> items;
    };
    std::vector<Cls> cls;
    std::vector<int> cls_of(G, -1);
    std::vector<Tile> tiles(G);
    int rc;
    for (int g = 0; g < G; ++g) {
        cfk_plan_s *p = m->plans[g];
        tiles[g] = plan_tile(p, CFK_PREC_MIXED);
        if (B[g] <= 0) continue;
        const Tile &t = tiles[g];
        const bool mx = t.pf == 1, split = mx && t.split;
        if (mx && (rc = ensure_mx_tables(p))) return rc;
        int c = 0;
        while (c < (int)cls.size() && !(cls[c].K == p->K && cls[c].KC == t.kc && cls[c].mx == mx && cls[c].split == split))
            ++c;
        if (c == (int)cls.size()) cls.push_back(Cls{p->K, t.kc, mx, split});
        cls_of[g] = c;
    }
    // work items of each class, heaviest first: cost of (group,row) from the plan's cost model, one
    // item per 64-distribution batch tile
    struct Row {
        int g, r;
        double w;
    };
    for (int c = 0; c < (int)cls.size(); ++c) {
        std::vector<Row> rows;
        for (int g = 0; g < G; ++g)
            if (cls_of[g] == c)
                for (int r = 0; r < m->plans[g]->N_Q; ++r) rows.push_back(Row{g, r, m->plans[g]->sym_row_work[r]});
        std::stable_sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.w > b.w; });
        for (const Row &rw : rows) {
            const int64_t tiles = (B[rw.g] + 63) / 64;
            for (int64_t t = 0; t < tiles; ++t) cls[c].items.push_back(make_int4(rw.g, rw.r, (int)t, 0));
            cls[c].work += rw.w * (double)tiles;
        }
    }
    // heaviest class first: it is launched first and so gets the GPU first
    std::stable_sort(cls.begin(), cls.end(), [](const Cls &a, const Cls &b) { return a.work > b.work; });
    std::vector<int4> items;
    std::vector<int> off(1, 0);
    for (const Cls &c : cls) {
        items.insert(items.end(), c.items.begin(), c.items.end());
        if (items.size() > (size_t)0x7fffffff) return set_error(CFK_EINVAL, "too many work items");
        off.push_back((int)items.size());
    }
    m->n_items = (int)items.size();
    if ((rc = m->items->upload(items.data(), items.size()))) return rc;
    m->f32.clear();
    m->packed.assign(G, 0);
    m->pack_kc.assign(G, 0);
    std::vector<MultiGroupDev> gd(G);
    for (int g = 0; g < G; ++g) {
        cfk_plan_s *p = m->plans[g];
        // (from the group's own tile, as its class was formed: cls has been re-sorted since); the
        // matrix-core kernel and VALU kernels with a full k3 chunk read packed FP32 rows (chunk-packed
        // for the matrix-core kernel's chunks)
        const bool mx = tiles[g].pf == 1;
        m->packed[g] = B[g] > 0 && (mx || tiles[g].kc >= p->K);
        m->pack_kc[g] = mx ? mx_chunk(p->K) : p->K;
        const int64_t n = (int64_t)p->N_Q * (m->packed[g] ? rl_of(p->K, m->pack_kc[g]) : p->K) * B[g];
        m->f32.emplace_back(new DevBuf<float>());
        if ((rc = m->f32.back()->alloc((size_t)std::max<int64_t>(n, 1)))) return rc;
        gd[g] = MultiGroupDev{m->f32.back()->p, Q[g], p->sym_slices.p, p->sym_row_ptr.p, p->sym_zq.p,
                              p->sym_zg32.p, p->sym_R.p, p->sym_R32.p, mx ? p->sym_Rmx.p : nullptr,
                              p->sym_row_mid.p, B[g], p->N_Q, p->K, p->accumulate};
    }
    if ((rc = m->groups->upload(gd.data(), gd.size()))) return rc;
    const int nc = (int)cls.size();
    m->launches.clear();
    for (int c = 0; c < nc; ++c) {
        if (!m->branch[c]) CFK_HIP(hipStreamCreateWithFlags(&m->branch[c], hipStreamNonBlocking));
        if (!m->ev[c]) CFK_HIP(hipEventCreateWithFlags(&m->ev[c], hipEventDisableTiming));
        m->launches.push_back(cfk_multi_s::Launch{cls[c].K, cls[c].KC, off[c], off[c + 1], cls[c].mx, cls[c].split});
    }
    if (!m->ev[MULTI_MAX_GROUPS]) CFK_HIP(hipEventCreateWithFlags(&m->ev[MULTI_MAX_GROUPS], hipEventDisableTiming));
    m->key_f.assign(f, f + G);
    m->key_Q.assign(Q, Q + G);
    m->key_B.assign(B, B + G);
    m->key_acc.resize(G);
    for (int g = 0; g < G; ++g) m->key_acc[g] = m->plans[g]->accumulate;
    m->prepared = true;
    return CFK_OK;
}

// The launch sequence on stream st: conversions, then a fork with one item kernel per K-class on its
// own branch stream, joined back into st.
static int multi_enqueue(cfk_multi_s *m, hipStream_t st) {
    const int G = (int)m->plans.size();
    int rc = CFK_OK;
    for (int g = 0; g < G && rc == CFK_OK; ++g) {
        const cfk_plan_s *p = m->plans[g];
        if (m->key_B[g] > 0)
            rc = m->packed[g] ? launch_pack_rows(m->key_f[g], m->f32[g]->p, p->N_Q, p->K, m->pack_kc[g], m->key_B[g],
                                                 false, st)
                              : launch_to_f32(m->key_f[g], m->f32[g]->p, (int64_t)p->N_Q * p->K * m->key_B[g], st);
    }
    const int nc = (int)m->launches.size();
    if (rc || nc == 0) return rc;
    CFK_HIP(hipEventRecord(m->ev[MULTI_MAX_GROUPS], st));
    for (int c = 0; c < nc; ++c) {
        const cfk_multi_s::Launch &l = m->launches[c];
        CFK_HIP(hipStreamWaitEvent(m->branch[c], m->ev[MULTI_MAX_GROUPS], 0));
        if ((rc = l.mx ? launch_multi_mx(m->groups->p, m->items->p, l.K, l.split, l.i0, l.i1, m->branch[c])
                       : launch_multi(m->groups->p, m->items->p, l.K, l.KC, l.i0, l.i1, m->branch[c])))
            return rc;
        CFK_HIP(hipEventRecord(m->ev[c], m->branch[c]));
        CFK_HIP(hipStreamWaitEvent(st, m->ev[c], 0));
    }
    return CFK_OK;
}

static int multi_capture(cfk_multi_s *m) {
    m->reset();
    CFK_HIP(hipStreamBeginCapture(m->cap, hipStreamCaptureModeThreadLocal));
    const int rc = multi_enqueue(m, m->cap);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(m->cap, &graph);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        return set_error(CFK_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    m->graph = graph;
    CFK_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    return CFK_OK;
}

extern "C" {

int cfk_multi_create(const cfk_plan *plans, int32_t n_groups, cfk_multi *out) {
    if (!out || !plans) return set_error(CFK_EINVAL, "null argument");
    *out = nullptr;
    if (n_groups < 1 || n_groups > MULTI_MAX_GROUPS)
        return set_error(CFK_EINVAL, "n_groups=%d outside 1..%d", n_groups, MULTI_MAX_GROUPS);
    int rc;
    if ((rc = check_device())) return rc;
    std::unique_ptr<cfk_multi_s> m(new cfk_multi_s());
    CFK_HIP(hipGetDevice(&m->device));
    for (int g = 0; g < n_groups; ++g) {
        cfk_plan_s *p = plans[g];
        if (!p) return set_error(CFK_EINVAL, "plan %d is null", g);
        if (p->device != m->device) return set_error(CFK_EINVAL, "plan %d lives on another device", g);
        if (!p->sym_ok) return set_error(CFK_EUNSUP, "plan %d: symmetrized tables unavailable: %s", g, p->sym_why.c_str());
        if (!multi_supports_k(p->K)) return set_error(CFK_EUNSUP, "plan %d: K=%d not in {6,8,10,12,14}", g, p->K);
        m->plans.push_back(p);
    }
    CFK_HIP(hipStreamCreateWithFlags(&m->cap, hipStreamNonBlocking));
    *out = m.release();
    return CFK_OK;
}

int cfk_multi_set_graph(cfk_multi m, int32_t use_graph) {
    if (!m) return set_error(CFK_EINVAL, "null multi plan");
    m->use_graph = use_graph != 0;
    return CFK_OK;
}

int cfk_multi_destroy(cfk_multi m) {
    delete m;
    return CFK_OK;
}

int cfk_multi_execute(cfk_multi m, double *const *Q_dev, const double *const *f_dev, const int64_t *B, void *stream) {
    if (!m || !Q_dev || !f_dev || !B) return set_error(CFK_EINVAL, "null argument");
    const int G = (int)m->plans.size();
    for (int g = 0; g < G; ++g) {
        if (B[g] < 0 || B[g] > max_batch(m->plans[g]->N_Q, m->plans[g]->K))
            return set_error(CFK_EINVAL, "B[%d] = %lld out of range", g, (long long)B[g]);
        if (B[g] > 0 && (!Q_dev[g] || !f_dev[g])) return set_error(CFK_EINVAL, "null buffer for group %d", g);
    }
    int dev = -1;
    CFK_HIP(hipGetDevice(&dev));
    if (dev != m->device) return set_error(CFK_EINVAL, "multi plan lives on device %d, current is %d", m->device, dev);
    const bool same = m->prepared && std::equal(m->key_f.begin(), m->key_f.end(), f_dev) &&
                      std::equal(m->key_Q.begin(), m->key_Q.end(), Q_dev) &&
                      std::equal(m->key_B.begin(), m->key_B.end(), B) &&
                      std::equal(m->key_acc.begin(), m->key_acc.end(), m->plans.begin(),
                                 [](int a, const cfk_plan_s *p) { return a == p->accumulate; });
    int rc;
    if (!same && (rc = multi_prepare(m, Q_dev, f_dev, B))) return rc;
    if (!m->launched) CFK_HIP(hipEventCreateWithFlags(&m->launched, hipEventDisableTiming));
    if (!m->use_graph || m->frozen) {
        if ((rc = multi_enqueue(m, (hipStream_t)stream))) return rc;
    } else {
        if (!m->exec && (rc = multi_capture(m))) return rc;
        CFK_HIP(hipGraphLaunch(m->exec, (hipStream_t)stream));
    }
    CFK_HIP(hipEventRecord(m->launched, (hipStream_t)stream));
    return CFK_OK;
}

}  // extern "C"
