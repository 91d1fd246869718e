// Serving loop: continuous batching of energy / force / Mahalanobis requests.
//
// Each request is one structure.  A worker thread owns NSLOT private contexts
// (each with its own stream, batch state and a copy of the resident model) and
// keeps up to NSLOT batches in flight: whenever a slot is free it admits the
// oldest waiting requests, as many as fit the per-batch limits and the HBM budget
// left by the batches already on the device, and launches
//   upload -> neighbour lists -> descriptors (+ dB/dr) -> prediction -> download
// on that slot's stream, so the work of consecutive batches overlaps on the GPU.
// Every result depends only on its own structure (per-structure sums in a fixed
// order), so a request returns the same numbers alone or in any mix.
// The reference evaluates one configuration at a time (md_mld_compute_energy /
// md_mld_compute_force, mld.F90:1579-1772; compute_mcd_dist_for_one_config,
// main_descriptors_analysis.F90:1315-1405); it has no server.
#include "mld_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

namespace {

constexpr int NSLOT = 2;

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    long long id = 0;
    int nat = 0, want_f = 0, want_d2 = 0;
    std::vector<double> pos, cell;  // pos [atom][xyz], cell column-major
    std::vector<int> itype;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    int status = MLD_OK;
    double E = 0;
    std::vector<double> F, d2;
    size_t bytes = 0;
};

struct Slot {
    mld_ctx* ctx = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Request*> reqs;
    size_t bytes = 0;
    int natoms = 0, nstruct = 0, want_f = 0, want_d2 = 0;
    long long npairs = 0;
    std::vector<int> atom_off;
    // pinned host staging and device buffers (grown geometrically)
    void* h_in = nullptr; size_t cap_h_in = 0;
    void* h_out = nullptr; size_t cap_h_out = 0;
    void* d_in = nullptr; size_t cap_d_in = 0;
    void* d_out = nullptr; size_t cap_d_out = 0;
    void* d_desc = nullptr; size_t cap_d_desc = 0;
};

}  // namespace

struct mld_server {
    mld_ctx* parent = nullptr;
    size_t budget = 0;
    int max_structs = 0, max_atoms = 0;
    size_t nn_est = 0;
    Slot slots[NSLOT];
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<long long, Request*> all;
    long long next_id = 1;
    bool stop = false;
    std::thread worker;
    std::string err;
    // statistics of completed requests
    std::vector<double> lat_ms;
    double t_first = -1, t_last = 0;
    long long atoms_done = 0, batches = 0, structs_in_batches = 0;
    int max_batch = 0;
    size_t inflight = 0, max_inflight = 0;
    bool hold = false;  // the queue front waits for the batch in flight to finish
};

namespace {

int grow_pinned(mld_ctx* c, void** p, size_t* cap, size_t need)
{
    if (need <= *cap) return MLD_OK;
    if (*p) MLD_HIP(c, hipHostFree(*p));
    *p = nullptr;
    size_t n = std::max(need, *cap * 2);
    MLD_HIP(c, hipHostMalloc(p, n, hipHostMallocDefault));
    *cap = n;
    return MLD_OK;
}

int grow_dev(mld_ctx* c, void** p, size_t* cap, size_t need)
{
    if (need <= *cap) return MLD_OK;
    return device_ensure(c, p, cap, std::max(need, *cap * 2));
}

// Device bytes a batch holds while it is in flight: inputs, pair list and per-pair geometry
// record, descriptors (and gradients), outputs.  Linear in its arguments, so the bytes of a batch
// are the sum of the bytes of its structures.
size_t batch_bytes(const mld_ctx* c, size_t natoms, size_t nstruct, size_t npairs, int want_f)
{
    const size_t esz = c->p.precision == MLD_PREC_FP32 ? 4 : 8;
    size_t b = natoms * (3 * sizeof(double) + sizeof(int)) + nstruct * 9 * sizeof(double);
    b += (natoms + 1) * sizeof(int) + npairs * (sizeof(int) + 4 * sizeof(double) + 24 * esz);
    b += natoms * (size_t)c->D * esz;
    if (want_f) b += (npairs + natoms) * 3 * (size_t)c->D * esz;
    b += (nstruct + 4 * natoms) * sizeof(double);
    return b;
}

// First phase of a batch held by slot s: upload and neighbour lists; fills act[i] with the
// device bytes request i needs at its actual pair count.
int launch_neighbours(mld_server* sv, Slot& s, std::vector<size_t>& act)
{
    mld_ctx* c = s.ctx;
    const int ns = (int)s.reqs.size();
    s.atom_off.assign(ns + 1, 0);
    s.want_f = s.want_d2 = 0;
    for (int i = 0; i < ns; ++i) {
        s.atom_off[i + 1] = s.atom_off[i] + s.reqs[i]->nat;
        s.want_f |= s.reqs[i]->want_f;
        s.want_d2 |= s.reqs[i]->want_d2;
    }
    const int na = s.atom_off[ns];
    s.natoms = na;
    s.nstruct = ns;
    // inputs: pos (3 na doubles), cells (9 ns doubles), itype (na ints)
    const size_t b_pos = 3 * (size_t)na * sizeof(double), b_cell = 9 * (size_t)ns * sizeof(double);
    const size_t b_it = (size_t)na * sizeof(int);
    int rc = grow_pinned(c, &s.h_in, &s.cap_h_in, b_pos + b_cell + b_it);
    if (!rc) rc = grow_dev(c, &s.d_in, &s.cap_d_in, b_pos + b_cell + b_it);
    if (rc) return rc;
    char* hp = (char*)s.h_in;
    for (int i = 0; i < ns; ++i) {
        const Request* r = s.reqs[i];
        std::copy(r->pos.begin(), r->pos.end(), (double*)hp + 3 * (size_t)s.atom_off[i]);
        std::copy(r->cell.begin(), r->cell.end(), (double*)(hp + b_pos) + 9 * (size_t)i);
        std::copy(r->itype.begin(), r->itype.end(), (int*)(hp + b_pos + b_cell) + s.atom_off[i]);
    }
    MLD_HIP(c, hipMemcpyAsync(s.d_in, s.h_in, b_pos + b_cell + b_it, hipMemcpyHostToDevice, c->stream));
    const double* d_pos = (const double*)s.d_in;
    const double* d_cell = (const double*)((char*)s.d_in + b_pos);
    const int* d_it = (const int*)((char*)s.d_in + b_pos + b_cell);
    (void)d_cell;
    // the cells and species are on the host already: the neighbour set-up reads nothing back
    // before the count, and the pair offsets come back with the pair total (one wait per batch)
    std::vector<int> off(na + 1);
    rc = batch_neighbours_host(c, ns, s.atom_off.data(), d_pos, (const double*)(hp + b_pos), d_it,
                               (const int*)(hp + b_pos + b_cell), off.data());
    if (rc) return rc;
    s.npairs = c->npairs;
    // pair counts per structure from the CSR offsets
    act.assign(ns, 0);
    for (int i = 0; i < ns; ++i)
        act[i] = batch_bytes(c, s.reqs[i]->nat, 1, (size_t)(off[s.atom_off[i + 1]] - off[s.atom_off[i]]),
                             s.reqs[i]->want_f);
    return MLD_OK;
}

// Second phase: descriptors, prediction and download on the slot's stream.
int launch_rest(mld_server* sv, Slot& s)
{
    (void)sv;
    mld_ctx* c = s.ctx;
    const int ns = s.nstruct, na = s.natoms;
    const long long P = s.npairs;
    int rc = MLD_OK;
    const size_t esz = c->p.precision == MLD_PREC_FP32 ? 4 : 8;
    const size_t b_e = (size_t)na * c->D * esz;
    const size_t b_g = s.want_f ? (size_t)(P + na) * 3 * c->D * esz : 0;
    rc = grow_dev(c, &s.d_desc, &s.cap_d_desc, b_e + b_g + 256);
    if (rc) return rc;
    char* dd = (char*)s.d_desc;
    rc = mld_batch_bso4_dev(c, s.want_f, dd, s.want_f ? dd + ((b_e + 255) / 256) * 256 : nullptr);
    if (rc) return rc;
    // outputs: E (ns), F (3 na), d2 (na)
    const size_t b_out = ((size_t)ns + 4 * (size_t)na) * sizeof(double);
    rc = grow_dev(c, &s.d_out, &s.cap_d_out, b_out);
    if (!rc) rc = grow_pinned(c, &s.h_out, &s.cap_h_out, b_out);
    if (rc) return rc;
    double* dE = (double*)s.d_out;
    double* dF = dE + ns;
    double* dQ = dF + 3 * (size_t)na;
    rc = mld_predict_dev(c, dE, s.want_f ? dF : nullptr, s.want_d2 ? dQ : nullptr);
    if (rc) return rc;
    MLD_HIP(c, hipMemcpyAsync(s.h_out, s.d_out, b_out, hipMemcpyDeviceToHost, c->stream));
    MLD_HIP(c, hipEventRecord(s.ev, c->stream));
    return MLD_OK;
}

// Results of a finished slot to its requests (caller holds the lock).
void complete(mld_server* sv, Slot& s, int status)
{
    const double t = now_s();
    const double* hE = (const double*)s.h_out;
    const double* hF = hE + s.nstruct;
    const double* hQ = hF + 3 * (size_t)s.natoms;
    for (int i = 0; i < s.nstruct; ++i) {
        Request* r = s.reqs[i];
        r->status = status;
        if (status == MLD_OK) {
            r->E = hE[i];
            if (r->want_f) r->F.assign(hF + 3 * (size_t)s.atom_off[i], hF + 3 * (size_t)s.atom_off[i + 1]);
            if (r->want_d2) r->d2.assign(hQ + s.atom_off[i], hQ + s.atom_off[i + 1]);
        }
        r->t_done = t;
        r->done = true;
        sv->lat_ms.push_back(1e3 * (t - r->t_submit));
        sv->atoms_done += r->nat;
        sv->t_last = t;
        r->pos.clear();
        r->pos.shrink_to_fit();
    }
    sv->inflight -= s.bytes;
    s.reqs.clear();
    s.busy = false;
    s.bytes = 0;
}

void worker_main(mld_server* sv)
{
    (void)hipSetDevice(sv->parent->p.device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool any_busy = false;
        for (Slot& s : sv->slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.ev);
            if (q == hipSuccess) complete(sv, s, MLD_OK);
            else if (q != hipErrorNotReady) complete(sv, s, MLD_E_DEVICE);
            else any_busy = true;
            if (!s.busy) sv->hold = false;
        }
        sv->cv_done.notify_all();
        if (sv->stop && sv->queue.empty() && !any_busy) break;
        // admit waiting requests into a free slot
        Slot* free_slot = nullptr;
        for (Slot& s : sv->slots)
            if (!s.busy) { free_slot = &s; break; }
        if (free_slot && !sv->queue.empty() && !sv->hold) {
            Slot& s = *free_slot;
            int na = 0;
            size_t bytes = 0;
            while (!sv->queue.empty()) {
                Request* r = sv->queue.front();
                const bool first = s.reqs.empty();
                if (!first && ((int)s.reqs.size() >= sv->max_structs || na + r->nat > sv->max_atoms)) break;
                if (sv->inflight + bytes + r->bytes > sv->budget && !(first && sv->inflight == 0)) break;
                s.reqs.push_back(r);
                na += r->nat;
                bytes += r->bytes;
                sv->queue.pop_front();
            }
            if (!s.reqs.empty()) {
                s.busy = true;
                s.bytes = bytes;
                sv->inflight += bytes;
                sv->batches += 1;
                sv->structs_in_batches += (long long)s.reqs.size();
                const double t = now_s();
                for (Request* r : s.reqs) r->t_start = t;
                // the admission above used the requests' estimated bytes; with the neighbour lists
                // built, the batch reserves its actual bytes and gives back the requests that no
                // longer fit (a batch keeps at least one request when nothing else is in flight)
                std::vector<size_t> act;
                lk.unlock();
                int rc = launch_neighbours(sv, s, act);
                lk.lock();
                if (!rc) {
                    const size_t other = sv->inflight - s.bytes;
                    size_t tot = 0;
                    size_t k = 0;
                    while (k < s.reqs.size()) {
                        if (other + tot + act[k] > sv->budget && !(k == 0 && other == 0)) break;
                        tot += act[k];
                        ++k;
                    }
                    if (k < s.reqs.size()) {
                        for (size_t i = s.reqs.size(); i-- > k;) sv->queue.push_front(s.reqs[i]);
                        sv->structs_in_batches -= (long long)(s.reqs.size() - k);
                        s.reqs.resize(k);
                    }
                    sv->inflight = other;
                    s.bytes = 0;
                    if (k == 0) {
                        // the batch in flight must finish first
                        s.busy = false;
                        sv->batches -= 1;
                        sv->hold = true;
                        continue;
                    }
                    sv->max_batch = std::max(sv->max_batch, (int)k);
                    if (act.size() != k) {  // rebuild the lists of the kept requests
                        lk.unlock();
                        rc = launch_neighbours(sv, s, act);
                        lk.lock();
                    }
                    s.bytes = tot;
                    sv->inflight += tot;
                    sv->max_inflight = std::max(sv->max_inflight, sv->inflight);
                    if (!rc) {
                        lk.unlock();
                        rc = launch_rest(sv, s);
                        lk.lock();
                    }
                }
                if (rc) {
                    sv->err = mld_last_error(s.ctx);
                    (void)hipStreamSynchronize(s.ctx->stream);
                    complete(sv, s, rc);
                    sv->cv_done.notify_all();
                }
                continue;
            }
        }
        // nothing to admit: wait for a request, or poll the device briefly
        if (any_busy)
            sv->cv_work.wait_for(lk, std::chrono::microseconds(50));
        else
            sv->cv_work.wait_for(lk, std::chrono::milliseconds(20));
    }
}

}  // namespace

extern "C" {

int mld_serve_create(mld_ctx* ctx, size_t hbm_budget_bytes, int max_batch_structs, int max_batch_atoms,
                     mld_server** out)
{
    if (!ctx || !out) return MLD_E_ARG;
    if (max_batch_structs < 1 || max_batch_atoms < 1 || hbm_budget_bytes == 0)
        return mld_fail(ctx, MLD_E_ARG, "mld_serve_create: limits must be positive");
    if (!ctx->predict_state) return mld_fail(ctx, MLD_E_STATE, "mld_serve_create: mld_model_set first");
    if (predict_order(ctx) == 3)
        return mld_fail(ctx, MLD_E_UNSUPPORTED, "mld_serve_create: kernel models (order 3) are not served");
    mld_server* sv = new mld_server();
    sv->parent = ctx;
    sv->budget = hbm_budget_bytes;
    sv->max_structs = max_batch_structs;
    sv->max_atoms = max_batch_atoms;
    // admission estimate of a request's neighbours per atom (number density 0.12 / A^3 inside
    // r_cut); a batch then reserves the bytes of its actual pair count (worker_main)
    const double rc = ctx->p.r_cut;
    sv->nn_est = (size_t)std::ceil(4.0 / 3.0 * 3.14159265358979323846 * rc * rc * rc * 0.12);
    for (Slot& s : sv->slots) {
        int rc2 = mld_create(&ctx->p, &s.ctx);
        if (rc2 == MLD_OK) rc2 = predict_clone(s.ctx, ctx);
        if (rc2 == MLD_OK && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) rc2 = MLD_E_DEVICE;
        if (rc2 != MLD_OK) {
            std::string m = s.ctx ? mld_last_error(s.ctx) : "context creation failed";
            for (Slot& t : sv->slots) {
                if (t.ev) (void)hipEventDestroy(t.ev);
                if (t.ctx) mld_destroy(t.ctx);
            }
            delete sv;
            return mld_fail(ctx, rc2, "mld_serve_create: " + m);
        }
    }
    sv->worker = std::thread(worker_main, sv);
    *out = sv;
    return MLD_OK;
}

int mld_serve_submit(mld_server* sv, int nat, const double* pos, const double* cell, const int* itype,
                     int want_forces, int want_d2, long long* id)
{
    if (!sv || !id) return MLD_E_ARG;
    mld_ctx* c = sv->parent;
    if (nat < 1 || !pos || !cell || !itype) return mld_fail(c, MLD_E_ARG, "mld_serve_submit: empty structure");
    for (int a = 0; a < nat; ++a)
        if (itype[a] < 0 || itype[a] >= c->nsp) return mld_fail(c, MLD_E_ARG, "mld_serve_submit: itype out of range");
    Request* r = new Request();
    r->nat = nat;
    r->want_f = want_forces ? 1 : 0;
    r->want_d2 = want_d2 ? 1 : 0;
    r->pos.resize(3 * (size_t)nat);
    for (int a = 0; a < nat; ++a)
        for (int x = 0; x < 3; ++x) r->pos[3 * (size_t)a + x] = pos[3 * (size_t)a + x];
    r->cell.assign(cell, cell + 9);
    r->itype.assign(itype, itype + nat);
    r->bytes = batch_bytes(c, (size_t)nat, 1, (size_t)nat * sv->nn_est, r->want_f);
    if (r->bytes > sv->budget) {
        delete r;
        return mld_fail(c, MLD_E_ARG, "mld_serve_submit: structure exceeds the HBM budget");
    }
    if (nat > sv->max_atoms) {
        delete r;
        return mld_fail(c, MLD_E_ARG, "mld_serve_submit: structure exceeds max_batch_atoms");
    }
    {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->stop) { delete r; return mld_fail(c, MLD_E_STATE, "mld_serve_submit: server stopping"); }
        r->id = sv->next_id++;
        r->t_submit = now_s();
        if (sv->t_first < 0) sv->t_first = r->t_submit;
        sv->queue.push_back(r);
        sv->all[r->id] = r;
        *id = r->id;
    }
    sv->cv_work.notify_one();
    return MLD_OK;
}

int mld_serve_poll(mld_server* sv, long long id, int wait, double* E, double* F, double* d2, double* latency_ms,
                   int* done)
{
    if (!sv || !done) return MLD_E_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return mld_fail(sv->parent, MLD_E_ARG, "mld_serve_poll: unknown request id");
    Request* r = it->second;
    if (wait) sv->cv_done.wait(lk, [&] { return r->done; });
    *done = r->done ? 1 : 0;
    if (!r->done) return MLD_OK;
    int st = r->status;
    if (st == MLD_OK) {
        if (E) *E = r->E;
        if (F && r->want_f) std::copy(r->F.begin(), r->F.end(), F);
        if (d2 && r->want_d2) std::copy(r->d2.begin(), r->d2.end(), d2);
    }
    if (latency_ms) *latency_ms = 1e3 * (r->t_done - r->t_submit);
    sv->all.erase(it);
    delete r;
    if (st != MLD_OK) return mld_fail(sv->parent, st, "mld_serve_poll: batch failed: " + sv->err);
    return MLD_OK;
}

int mld_serve_stats(mld_server* sv, double* stats)
{
    if (!sv || !stats) return MLD_E_ARG;
    std::lock_guard<std::mutex> g(sv->mu);
    std::vector<double> l = sv->lat_ms;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        size_t k = (size_t)std::ceil(q * l.size()) - 1;  // nearest-rank percentile
        return l[std::min(k, l.size() - 1)];
    };
    double mean = 0;
    for (double v : l) mean += v;
    mean = l.empty() ? 0 : mean / l.size();
    const double span = (sv->t_first >= 0 && sv->t_last > sv->t_first) ? sv->t_last - sv->t_first : 0;
    stats[0] = (double)l.size();
    stats[1] = pct(0.50);
    stats[2] = pct(0.99);
    stats[3] = mean;
    stats[4] = span > 0 ? l.size() / span : 0;
    stats[5] = span > 0 ? sv->atoms_done / span : 0;
    stats[6] = sv->max_batch;
    stats[7] = sv->batches ? (double)sv->structs_in_batches / sv->batches : 0;
    stats[8] = (double)sv->max_inflight;
    stats[9] = l.empty() ? 0 : l.back();
    return MLD_OK;
}

int mld_serve_destroy(mld_server* sv)
{
    if (!sv) return MLD_OK;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    for (Slot& s : sv->slots) {
        if (s.ev) (void)hipEventDestroy(s.ev);
        if (s.h_in) (void)hipHostFree(s.h_in);
        if (s.h_out) (void)hipHostFree(s.h_out);
        void* dv[] = {s.d_in, s.d_out, s.d_desc};
        for (void* p : dv)
            if (p) (void)hipFree(p);
        if (s.ctx) mld_destroy(s.ctx);
    }
    for (auto& kv : sv->all) delete kv.second;
    delete sv;
    return MLD_OK;
}

}  // extern "C"
