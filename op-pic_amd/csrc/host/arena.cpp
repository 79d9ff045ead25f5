// Batched ensembles and continuous-batching serving for C++ hosts: a resident arena of FEM-PIC
// instances stepped together, written only against the oppx C-ABI (no torch, no Python).
//
// An instance is admitted into first-fit ranges of one concatenated mesh (cells, nodes, overlay words)
// and appended to one shared particle set, its cell, node and p2c indices offset by its ranges' bases
// (the layout of oppx_batch_inst). Every step is one launch sequence over all resident instances:
// oppx_fempic_batch_step_fused (push, direct hop, walk), the Sort_All finalize (oppx_sort_by_cell,
// stable by concatenated cell, so each instance's particles stay contiguous and in the order they
// would have alone) and the FP32 compensated deposit. Retiring an instance finds its particle range
// by a lower-bound search in the sorted set, copies its state out, marks its particles removed (they
// leave the set at the next sort) and frees its ranges, so another instance can be admitted between
// any two steps. Every instance's results equal the instance stepped alone with the same finalize
// and deposit. OP-PIC has no counterpart: it runs one instance per process.

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "oppx.h"

namespace {

// First-fit allocator of index ranges in [0, cap).
class RangeAlloc {
public:
    explicit RangeAlloc(int64_t cap = 0) : cap_(cap) {
        if (cap > 0) free_.push_back({0, cap});
    }
    bool fits(int64_t n) const {
        for (const auto& f : free_)
            if (f.second >= n) return true;
        return false;
    }
    int64_t alloc(int64_t n) {
        for (size_t k = 0; k < free_.size(); k++) {
            if (free_[k].second < n) continue;
            const int64_t s = free_[k].first;
            if (free_[k].second == n) free_.erase(free_.begin() + (long)k);
            else free_[k] = {s + n, free_[k].second - n};
            used_[s] = n;
            return s;
        }
        return -1;
    }
    void release(int64_t s) {
        auto it = used_.find(s);
        if (it == used_.end()) return;
        free_.push_back({s, it->second});
        used_.erase(it);
        std::sort(free_.begin(), free_.end());
        std::vector<std::pair<int64_t, int64_t>> merged;
        for (const auto& f : free_) {
            if (!merged.empty() && merged.back().first + merged.back().second == f.first) merged.back().second += f.second;
            else merged.push_back(f);
        }
        free_.swap(merged);
    }
    int64_t extent() const { return used_.empty() ? 0 : used_.rbegin()->first + used_.rbegin()->second; }

private:
    int64_t cap_;
    std::vector<std::pair<int64_t, int64_t>> free_;  // sorted, disjoint (start, length)
    std::map<int64_t, int64_t> used_;                // start -> length
};

struct Resident {
    bool used = false;
    int64_t n_cells = 0, n_nodes = 0, n_part = 0, overlay_words = 0;
    int64_t cb = -1, nb = -1, ob = -1;
};

constexpr int32_t kMaxCell = OPPX_MAX_CELL_INDEX;

}  // namespace

struct oppx_arena {
    int64_t C = 0, N = 0, P = 0, O = 0;
    int32_t max_insts = 0;
    void* stream = nullptr;
    // device state
    double* det = nullptr;
    oppx_tet_aux* aux = nullptr;
    int32_t* c2n = nullptr;
    double* ef = nullptr;
    int32_t* cell_inst = nullptr;
    int32_t* overlay = nullptr;
    int32_t *node_start = nullptr, *node_end = nullptr, *node_ref = nullptr;
    float *charge_f32 = nullptr, *cell_partials = nullptr;
    oppx_batch_inst* inst_table = nullptr;
    double *pos = nullptr, *vel = nullptr, *lc = nullptr;
    int32_t* p2c = nullptr;
    double *pos2 = nullptr, *vel2 = nullptr, *lc2 = nullptr;
    int32_t* p2c2 = nullptr;
    oppx_move_counters* counters = nullptr;
    uint64_t* hole_bits = nullptr;
    void *work = nullptr, *sort_scratch = nullptr, *dep_scratch = nullptr;
    size_t work_bytes = 0, sort_bytes = 0, dep_bytes = 0;
    int32_t* keys = nullptr;   // [2] lower-bound keys of a retirement
    int64_t* bounds = nullptr; // [2]
    // host bookkeeping
    RangeAlloc cells, nodes, ovl;
    std::vector<Resident> res;
    std::vector<int32_t> slots_free;
    int64_t size = 0;     // particles in the set (retired ones included until the next sort)
    bool sorted = true;   // the set is ordered by concatenated cell (true after a step, false after an admission)
    std::vector<std::pair<int64_t, int64_t>> retired;  // (first, count) to mark removed before the next step
    std::vector<void*> owned;

    template <typename T>
    int grab(T*& p, int64_t count) {
        void* v = nullptr;
        const int rc = oppx_device_malloc(&v, (size_t)(count > 0 ? count : 1) * sizeof(T));
        if (rc != OPPX_OK) return rc;
        owned.push_back(v);
        p = static_cast<T*>(v);
        return oppx_memset_zero(v, (size_t)(count > 0 ? count : 1) * sizeof(T), stream);
    }
    int grab_bytes(void*& p, size_t bytes) {
        const int rc = oppx_device_malloc(&p, bytes > 0 ? bytes : 8);
        if (rc != OPPX_OK) return rc;
        owned.push_back(p);
        return oppx_memset_zero(p, bytes > 0 ? bytes : 8, stream);
    }
    ~oppx_arena() {
        for (void* p : owned) oppx_device_free(p);
    }
};

namespace {

int fill_i32(int32_t* dst, int64_t n, int32_t v, void* stream) {
    return oppx_fill_pattern(dst, n, 1, 4, &v, 0, stream);
}

int dat_descs(oppx_arena* a, oppx_dat_desc in[4], oppx_dat_desc out[4]) {
    in[0] = {a->pos, 3, 8};
    in[1] = {a->vel, 3, 8};
    in[2] = {a->lc, 4, 8};
    in[3] = {a->p2c, 1, 4};
    out[0] = {a->pos2, 3, 8};
    out[1] = {a->vel2, 3, 8};
    out[2] = {a->lc2, 4, 8};
    out[3] = {a->p2c2, 1, 4};
    return OPPX_OK;
}

}  // namespace

extern "C" {

int oppx_arena_create(int64_t max_cells, int64_t max_nodes, int64_t max_particles, int64_t max_overlay,
                      int32_t max_insts, void* stream, oppx_arena** out) {
    if (!out) return OPPX_ERR_ARG;
    *out = nullptr;
    if (max_cells <= 0 || max_cells >= kMaxCell || max_nodes <= 0 || max_nodes >= INT_MAX || max_particles <= 0 ||
        max_particles >= INT_MAX || max_overlay < 0 || max_insts <= 0 || 4 * max_cells >= (int64_t)INT_MAX)
        return OPPX_ERR_ARG;
    oppx_arena* a = new (std::nothrow) oppx_arena();
    if (!a) return OPPX_ERR_NOMEM;
    a->C = max_cells;
    a->N = max_nodes;
    a->P = max_particles;
    a->O = max_overlay > 0 ? max_overlay : 1;
    a->max_insts = max_insts;
    a->stream = stream;
    const int64_t C = a->C, N = a->N, P = a->P;
    int rc = OPPX_OK;
#define OPPX_GRAB(ptr, n) \
    if (rc == OPPX_OK) rc = a->grab(ptr, n)
    OPPX_GRAB(a->det, 16 * C);
    OPPX_GRAB(a->aux, C);
    OPPX_GRAB(a->c2n, 4 * C);
    OPPX_GRAB(a->ef, 3 * C);
    OPPX_GRAB(a->cell_inst, C);
    OPPX_GRAB(a->overlay, a->O);
    OPPX_GRAB(a->node_start, N);
    OPPX_GRAB(a->node_end, N);
    OPPX_GRAB(a->node_ref, 4 * C);
    OPPX_GRAB(a->charge_f32, N);
    OPPX_GRAB(a->cell_partials, 4 * C);
    OPPX_GRAB(a->inst_table, max_insts);
    OPPX_GRAB(a->pos, 3 * P);
    OPPX_GRAB(a->vel, 3 * P);
    OPPX_GRAB(a->lc, 4 * P);
    OPPX_GRAB(a->p2c, P);
    OPPX_GRAB(a->pos2, 3 * P);
    OPPX_GRAB(a->vel2, 3 * P);
    OPPX_GRAB(a->lc2, 4 * P);
    OPPX_GRAB(a->p2c2, P);
    OPPX_GRAB(a->counters, 1);
    OPPX_GRAB(a->hole_bits, oppx_hole_bits_words(P));
    OPPX_GRAB(a->keys, 2);
    OPPX_GRAB(a->bounds, 2);
#undef OPPX_GRAB
    a->work_bytes = oppx_fempic_work_bytes(P);
    a->sort_bytes = oppx_sort_scratch_bytes(P);
    a->dep_bytes = oppx_fempic_deposit_scratch_bytes(P);
    if (rc == OPPX_OK) rc = a->grab_bytes(a->work, a->work_bytes);
    if (rc == OPPX_OK) rc = a->grab_bytes(a->sort_scratch, a->sort_bytes);
    if (rc == OPPX_OK) rc = a->grab_bytes(a->dep_scratch, a->dep_bytes);
    if (rc == OPPX_OK) rc = fill_i32(a->cell_inst, C, -1, stream);
    if (rc == OPPX_OK) rc = fill_i32(a->p2c, P, kMaxCell, stream);
    if (rc == OPPX_OK) rc = oppx_stream_sync(stream);
    if (rc != OPPX_OK) {
        delete a;
        return rc;
    }
    a->cells = RangeAlloc(C);
    a->nodes = RangeAlloc(N);
    a->ovl = RangeAlloc(a->O);
    a->res.assign((size_t)max_insts, Resident{});
    for (int32_t s = max_insts - 1; s >= 0; s--) a->slots_free.push_back(s);
    *out = a;
    return OPPX_OK;
}

int oppx_arena_destroy(oppx_arena* a) {
    if (!a) return OPPX_OK;
    oppx_stream_sync(a->stream);
    delete a;
    return OPPX_OK;
}

int oppx_arena_admit(oppx_arena* a, const oppx_instance* in, int32_t* slot_out) {
    if (!a || !in || !slot_out) return OPPX_ERR_ARG;
    *slot_out = -1;
    const int64_t nc = in->n_cells, nn = in->n_nodes, n = in->n_part;
    const bool dh = in->overlay != nullptr;
    const int64_t words = dh ? in->dh_dims[0] * in->dh_dims[1] * in->dh_dims[2] : 0;
    if (nc <= 0 || nn <= 0 || n < 0 || !in->c_det || !in->c_vol || !in->c2c || !in->c2n || !in->c_ef ||
        (n > 0 && (!in->pos || !in->vel || !in->p2c)) || (dh && (words <= 0 || !(in->dh_inv_gs > 0))) ||
        !(in->mass != 0.0))
        return OPPX_ERR_ARG;
    // validate the instance's indices before anything is copied
    for (int64_t e = 0; e < 4 * nc; e++) {
        if (in->c2c[e] < -1 || in->c2c[e] >= nc) return OPPX_ERR_ARG;
        if (in->c2n[e] < 0 || in->c2n[e] >= nn) return OPPX_ERR_ARG;
    }
    for (int64_t i = 0; i < n; i++)
        if (in->p2c[i] < 0 || in->p2c[i] >= nc) return OPPX_ERR_ARG;
    for (int64_t w = 0; w < words; w++)
        if (in->overlay[w] != kMaxCell && (in->overlay[w] < 0 || in->overlay[w] >= nc)) return OPPX_ERR_ARG;
    // room now? (OPPX_ERR_NOMEM: not at the moment; retry after a retirement)
    if (a->slots_free.empty() || a->size + n > a->P || !a->cells.fits(nc) || !a->nodes.fits(nn) ||
        (dh && !a->ovl.fits(words)))
        return OPPX_ERR_NOMEM;

    // host-side records, indices rebased to the ranges
    const int64_t cb = a->cells.alloc(nc), nb = a->nodes.alloc(nn), ob = dh ? a->ovl.alloc(words) : -1;
    const int32_t slot = a->slots_free.back();
    a->slots_free.pop_back();
    auto undo = [&](int code) {  // a failed admission leaves the arena as it was
        a->cells.release(cb);
        a->nodes.release(nb);
        if (ob >= 0) a->ovl.release(ob);
        a->slots_free.push_back(slot);
        return code;
    };
    std::vector<oppx_tet_aux> aux((size_t)nc);
    std::vector<int32_t> c2n((size_t)(4 * nc)), nstart((size_t)nn + 1), nref((size_t)(4 * nc)), nend((size_t)nn);
    for (int64_t c = 0; c < nc; c++) {
        oppx_tet_aux& r = aux[(size_t)c];
        std::memset(&r, 0, sizeof(r));
        r.c2 = (1.0 / 6.0) / in->c_vol[c];  // the division oppx_tet_aux_build performs
        for (int k = 0; k < 4; k++) {
            const int32_t v = in->c2c[4 * c + k];
            r.c2c[k] = v >= 0 ? (int32_t)(v + cb) : v;
            c2n[(size_t)(4 * c + k)] = (int32_t)(in->c2n[4 * c + k] + nb);
        }
    }
    int rc = oppx_node_cell_csr(nc, nn, in->c2n, nstart.data(), nref.data());
    if (rc != OPPX_OK) return undo(rc);
    for (int64_t v = 0; v < nn; v++) {
        nend[(size_t)v] = (int32_t)(nstart[(size_t)v + 1] + 4 * cb);
        nstart[(size_t)v] = (int32_t)(nstart[(size_t)v] + 4 * cb);
    }
    for (auto& e : nref) e = (int32_t)(e + 4 * cb);
    oppx_batch_inst rec;
    std::memset(&rec, 0, sizeof(rec));
    rec.cell_base = cb;
    rec.n_cells = nc;
    rec.node_base = nb;
    rec.n_nodes = nn;
    if (dh) {
        std::vector<int32_t> ov((size_t)words);
        for (int64_t w = 0; w < words; w++) ov[(size_t)w] = in->overlay[w] == kMaxCell ? kMaxCell : (int32_t)(in->overlay[w] + cb);
        rc = oppx_memcpy_htod(a->overlay + ob, ov.data(), (size_t)words * 4, a->stream);
        if (rc != OPPX_OK) return undo(rc);
        rec.dh_cell = a->overlay + ob;
        for (int d = 0; d < 3; d++) {
            rec.dims[d] = in->dh_dims[d];
            rec.min[d] = in->dh_min[d];
        }
        rec.inv_gs = in->dh_inv_gs;
    }
    rec.dt = in->dt;
    rec.k = in->charge / in->mass * in->dt;

    void* s = a->stream;
    auto htod = [&](void* dst, const void* src, size_t bytes) {
        if (rc == OPPX_OK) rc = oppx_memcpy_htod(dst, src, bytes, s);
    };
    htod(a->det + 16 * cb, in->c_det, (size_t)nc * 16 * 8);
    htod(a->aux + cb, aux.data(), (size_t)nc * sizeof(oppx_tet_aux));
    htod(a->c2n + 4 * cb, c2n.data(), (size_t)nc * 16);
    htod(a->ef + 3 * cb, in->c_ef, (size_t)nc * 24);
    htod(a->node_start + nb, nstart.data(), (size_t)nn * 4);
    htod(a->node_end + nb, nend.data(), (size_t)nn * 4);
    htod(a->node_ref + 4 * cb, nref.data(), (size_t)nc * 16);
    htod(a->inst_table + slot, &rec, sizeof(rec));
    if (rc == OPPX_OK) rc = fill_i32(a->cell_inst + cb, nc, slot, s);
    if (n > 0) {
        std::vector<double> comp((size_t)n);
        for (int d = 0; d < 3; d++) {  // AoS host records -> SoA components at [size, size + n)
            for (int64_t i = 0; i < n; i++) comp[(size_t)i] = in->pos[3 * i + d];
            htod(a->pos + d * a->P + a->size, comp.data(), (size_t)n * 8);
            for (int64_t i = 0; i < n; i++) comp[(size_t)i] = in->vel[3 * i + d];
            htod(a->vel + d * a->P + a->size, comp.data(), (size_t)n * 8);
        }
        std::vector<int32_t> p((size_t)n);
        for (int64_t i = 0; i < n; i++) p[(size_t)i] = (int32_t)(in->p2c[i] + cb);
        htod(a->p2c + a->size, p.data(), (size_t)n * 4);
        const double zero = 0.0;
        if (rc == OPPX_OK) rc = oppx_fill_pattern(a->lc + a->size, n, 4, 8, &zero, a->P, s);
    }
    if (rc == OPPX_OK) rc = oppx_stream_sync(s);
    if (rc != OPPX_OK) return undo(rc);
    Resident& r = a->res[(size_t)slot];
    r.used = true;
    r.n_cells = nc;
    r.n_nodes = nn;
    r.n_part = n;
    r.overlay_words = words;
    r.cb = cb;
    r.nb = nb;
    r.ob = ob;
    a->size += n;
    if (n > 0) a->sorted = false;
    *slot_out = slot;
    return OPPX_OK;
}

int oppx_arena_step(oppx_arena* a) {
    if (!a) return OPPX_ERR_ARG;
    void* s = a->stream;
    const int64_t n_cells = std::max<int64_t>(1, a->cells.extent());
    const int64_t n_nodes = std::max<int64_t>(1, a->nodes.extent());
    int32_t n_insts = 1;
    for (int32_t k = 0; k < a->max_insts; k++)
        if (a->res[(size_t)k].used) n_insts = k + 1;
    const oppx_tet_mesh mesh{n_cells, n_nodes, a->det, a->aux, a->c2n, a->ef};
    const oppx_fempic_parts parts{a->P, a->pos, a->vel, a->lc, a->p2c};
    int rc = OPPX_OK;
    for (const auto& rr : a->retired)  // retired instances' particles: removed, dropped by this step's sort
        if (rc == OPPX_OK) rc = fill_i32(a->p2c + rr.first, rr.second, kMaxCell, s);
    a->retired.clear();
    if (rc == OPPX_OK) rc = oppx_fempic_step_begin(a->counters, nullptr, 0, s);
    const int32_t live = (int32_t)a->size;
    if (rc == OPPX_OK) rc = oppx_memcpy_htod(&a->counters->size, &live, 4, s);
    if (rc == OPPX_OK && a->size > 0)
        rc = oppx_fempic_batch_step_fused(a->size, &parts, &mesh, a->cell_inst, a->inst_table, n_insts, a->counters,
                                          a->hole_bits, a->work, a->work_bytes, nullptr, nullptr, s);
    oppx_dat_desc in[4], out[4];
    dat_descs(a, in, out);
    if (rc == OPPX_OK)
        rc = oppx_sort_by_cell(a->size, a->P, a->counters, n_cells, a->p2c, in, out, 4, nullptr, 1, a->sort_scratch,
                               a->sort_bytes, s);
    if (rc != OPPX_OK) return rc;
    std::swap(a->pos, a->pos2);
    std::swap(a->vel, a->vel2);
    std::swap(a->lc, a->lc2);
    std::swap(a->p2c, a->p2c2);
    rc = oppx_fempic_deposit_sorted_f32(a->size, a->P, a->counters, a->p2c, a->lc, n_cells, a->node_start, a->node_end,
                                        a->node_ref, n_nodes, a->cell_partials, nullptr, a->charge_f32,
                                        a->dep_scratch, a->dep_bytes, s);
    int32_t after = 0;
    if (rc == OPPX_OK) rc = oppx_memcpy_dtoh(&after, &a->counters->size, 4, s);  // waits for the step
    if (rc != OPPX_OK) return rc;
    a->size = after;
    a->sorted = true;
    return OPPX_OK;
}

int oppx_arena_retire(oppx_arena* a, int32_t slot, int64_t cap, int64_t* n_out, double* pos, double* vel, double* lc,
                      int32_t* p2c, float* charge_f32) {
    if (!a || !n_out || slot < 0 || slot >= a->max_insts || !a->res[(size_t)slot].used) return OPPX_ERR_ARG;
    // the range lookup needs the set in cell order: a step must separate an admission from a retirement
    if (!a->sorted) return OPPX_ERR_STATE;
    Resident& r = a->res[(size_t)slot];
    void* s = a->stream;
    const int32_t keys[2] = {(int32_t)r.cb, (int32_t)(r.cb + r.n_cells)};
    int rc = oppx_memcpy_htod(a->keys, keys, sizeof(keys), s);
    if (rc == OPPX_OK) rc = oppx_lower_bound(a->size, a->counters, a->p2c, a->keys, 2, a->bounds, s);
    int64_t b[2] = {0, 0};
    if (rc == OPPX_OK) rc = oppx_memcpy_dtoh(b, a->bounds, sizeof(b), s);
    if (rc != OPPX_OK) return rc;
    const int64_t lo = b[0], k = b[1] - b[0];
    *n_out = k;
    if (k < 0 || k > cap) return OPPX_ERR_ARG;  // buffers too small: the instance stays resident
    if (k > 0) {
        std::vector<double> comp((size_t)k);
        auto aos = [&](const double* soa, int dim, double* dst) {
            for (int d = 0; d < dim && rc == OPPX_OK; d++) {
                rc = oppx_memcpy_dtoh(comp.data(), soa + d * a->P + lo, (size_t)k * 8, s);
                if (dst)
                    for (int64_t i = 0; i < k; i++) dst[dim * i + d] = comp[(size_t)i];
            }
        };
        if (pos) aos(a->pos, 3, pos);
        if (vel) aos(a->vel, 3, vel);
        if (lc) aos(a->lc, 4, lc);
        if (p2c && rc == OPPX_OK) {
            rc = oppx_memcpy_dtoh(p2c, a->p2c + lo, (size_t)k * 4, s);
            for (int64_t i = 0; i < k; i++) p2c[i] = (int32_t)(p2c[i] - r.cb);
        }
    }
    if (charge_f32 && rc == OPPX_OK) rc = oppx_memcpy_dtoh(charge_f32, a->charge_f32 + r.nb, (size_t)r.n_nodes * 4, s);
    // its particles are marked removed at the start of the next step (so that later retirements still
    // search a sorted set) and leave it at that step's sort; its node rows become empty; its ranges are free
    if (rc == OPPX_OK && k > 0) a->retired.push_back({lo, k});
    if (rc == OPPX_OK) rc = fill_i32(a->node_start + r.nb, r.n_nodes, 0, s);
    if (rc == OPPX_OK) rc = fill_i32(a->node_end + r.nb, r.n_nodes, 0, s);
    if (rc == OPPX_OK) rc = oppx_stream_sync(s);
    if (rc != OPPX_OK) return rc;
    a->cells.release(r.cb);
    a->nodes.release(r.nb);
    if (r.ob >= 0) a->ovl.release(r.ob);
    r = Resident{};
    a->slots_free.push_back(slot);
    return OPPX_OK;
}

int oppx_arena_counters(const oppx_arena* a, oppx_move_counters* out) {
    if (!a || !out) return OPPX_ERR_ARG;
    return oppx_memcpy_dtoh(out, a->counters, sizeof(*out), a->stream);
}

}  // extern "C"
