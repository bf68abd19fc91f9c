// Batched uneven-size execution: several plans (different L and K) in one persistent kernel launch,
// captured once into a hipGraph and replayed (cfk_multi_* in include/cfk.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace cfk {
int launch_multi(const MultiDesc &md, int grid, hipStream_t st);
int launch_to_f32(const double *in, float *out, int64_t n, hipStream_t st);
int multi_blocks_per_cu();
bool multi_supports_k(int K);
}  // namespace cfk

using namespace cfk;

struct cfk_multi_s {
    std::vector<cfk_plan_s *> plans;
    int device = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the pointer/size set the current graph was captured for
    std::vector<const double *> key_f;
    std::vector<double *> key_Q;
    std::vector<int64_t> key_B;
    DevBuf<int4> items;
    DevBuf<int> counter;
    std::vector<std::unique_ptr<DevBuf<float>>> f32;
    int n_items = 0;
    ~cfk_multi_s() { reset(); if (cap) (void)hipStreamDestroy(cap); }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

static int multi_capture(cfk_multi_s *m, double *const *Q, const double *const *f, const int64_t *B) {
    const int G = (int)m->plans.size();
    m->reset();
    // work items, heaviest first: cost of (group,row) from the plan's cost model, one item per batch tile
    struct Row {
        int g, r;
        double w;
    };
    std::vector<Row> rows;
    for (int g = 0; g < G; ++g)
        if (B[g] > 0)
            for (int r = 0; r < m->plans[g]->N_Q; ++r) rows.push_back(Row{g, r, m->plans[g]->sym_row_work[r]});
    std::stable_sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.w > b.w; });
    std::vector<int4> items;
    for (const Row &rw : rows) {
        const int64_t tiles = (B[rw.g] + 63) / 64;
        for (int64_t t = 0; t < tiles; ++t) items.push_back(make_int4(rw.g, rw.r, (int)t, 0));
    }
    if (items.size() > (size_t)0x7fffffff) return set_error(CFK_EINVAL, "too many work items");
    m->n_items = (int)items.size();
    int rc;
    if ((rc = m->items.upload(items.data(), items.size()))) return rc;
    if ((rc = m->counter.ensure(1))) return rc;
    m->f32.clear();
    MultiDesc md{};
    for (int g = 0; g < G; ++g) {
        cfk_plan_s *p = m->plans[g];
        const int64_t n = (int64_t)p->N_Q * p->K * B[g];
        m->f32.emplace_back(new DevBuf<float>());
        if ((rc = m->f32.back()->alloc((size_t)std::max<int64_t>(n, 1)))) return rc;
        md.g[g] = MultiGroupDev{m->f32.back()->p, Q[g], p->sym_slices.p, p->sym_row_ptr.p, p->sym_zq.p,
                                p->sym_zg32.p, p->sym_R.p, p->sym_R32.p, B[g], p->N_Q, p->K};
    }
    md.items = m->items.p;
    md.n_items = m->n_items;
    md.counter = m->counter.p;
    int ncu = 0;
    CFK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device));
    const int grid = std::max(1, std::min(m->n_items, ncu * multi_blocks_per_cu()));
    CFK_HIP(hipStreamBeginCapture(m->cap, hipStreamCaptureModeThreadLocal));
    rc = CFK_OK;
    for (int g = 0; g < G && rc == CFK_OK; ++g)
        if (B[g] > 0) rc = launch_to_f32(f[g], m->f32[g]->p, (int64_t)m->plans[g]->N_Q * m->plans[g]->K * B[g], m->cap);
    if (rc == CFK_OK && m->n_items > 0) rc = launch_multi(md, grid, m->cap);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(m->cap, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_error(CFK_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    m->graph = graph;
    CFK_HIP(hipGraphInstantiate(&m->exec, m->graph, nullptr, nullptr, 0));
    m->key_f.assign(f, f + G);
    m->key_Q.assign(Q, Q + G);
    m->key_B.assign(B, B + G);
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

int cfk_multi_destroy(cfk_multi m) {
    delete m;
    return CFK_OK;
}

int cfk_multi_execute(cfk_multi m, double *const *Q_dev, const double *const *f_dev, const int64_t *B, void *stream) {
    if (!m || !Q_dev || !f_dev || !B) return set_error(CFK_EINVAL, "null argument");
    const int G = (int)m->plans.size();
    for (int g = 0; g < G; ++g) {
        if (B[g] < 0) return set_error(CFK_EINVAL, "B[%d] < 0", g);
        if (B[g] > 0 && (!Q_dev[g] || !f_dev[g])) return set_error(CFK_EINVAL, "null buffer for group %d", g);
    }
    int dev = -1;
    CFK_HIP(hipGetDevice(&dev));
    if (dev != m->device) return set_error(CFK_EINVAL, "multi plan lives on device %d, current is %d", m->device, dev);
    const bool same = m->exec && std::equal(m->key_f.begin(), m->key_f.end(), f_dev) &&
                      std::equal(m->key_Q.begin(), m->key_Q.end(), Q_dev) &&
                      std::equal(m->key_B.begin(), m->key_B.end(), B);
    int rc;
    if (!same && (rc = multi_capture(m, Q_dev, f_dev, B))) return rc;
    CFK_HIP(hipGraphLaunch(m->exec, (hipStream_t)stream));
    return CFK_OK;
}

}  // extern "C"
