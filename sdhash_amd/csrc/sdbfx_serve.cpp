// Query serving over a resident reference set: continuous batching on two
// HIP streams.  Built on the engine's C ABI only.
//
// Requests are queued by sdbfx_serve_submit.  A digest thread takes whatever
// is queued (in arrival order, up to batch_bytes of input and batch_max
// queries; a larger query goes alone) and digests it as one batch on its own
// context; the compare thread compares each digested batch with the
// reference set on a second context.  While batch k is compared, batch k+1
// is digested, and requests that arrive meanwhile form batch k+2.  At most
// two batches are past the queue at a time.
//
// Admission by memory: each of the two contexts gets half of the server's
// memory budget for its scratch (staging, work lists, partial results).  A
// batch whose digesting or comparison would exceed it is split: its later
// half goes back to the front of the queue (deferred, still in arrival order)
// and the earlier half is tried again, after the context's cached scratch is
// freed.  A single query that does not fit on its own is refused with status
// SDBFX_EBUDGET.  The query sets themselves (256 bytes per filter, about 1/8
// of the input in stream mode at most) are not counted.
//
// Autotuning: the two comparison kernels give identical scores at different
// speeds depending on the shape of the batch (query filters per reference
// digest) and on how many filter pairs come close to their cutoff.  The first
// batch of each size class (log2 of its filter count) is compared with both
// against a sample of the reference set (its first whole digests, at least
// 65,536 filters, copied once when the server starts), and the faster one is
// used for that class from then on.  When the engine's cost model puts the
// MFMA kernel under half the VALU kernel's cost, the trial is skipped.
//
// The reference counterpart is sdhash-server's handler (sdhash-srv-main.cc:
// 224-402), which hashes and compares each request on its own.
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sdbfx.h"

namespace {

// batches of at most this many query filters are timed with both kernels
constexpr uint64_t SERVE_TUNE_SAMPLE_FILTERS = 65536;

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    uint64_t ticket = 0;
    std::vector<uint8_t> data;
    std::string name;
    uint32_t dd = 0;
    uint64_t t_submit = 0, t_done = 0;
    bool done = false;
    int32_t status = 0;
    std::vector<int32_t> scores;
};

struct Batch {
    std::vector<std::shared_ptr<Request>> reqs;
    sdbfx_set* set = nullptr;      // digests of the requests with status 0, in order
    std::vector<int32_t> status;
};

}  // namespace

struct sdbfx_server {
    const sdbfx_set* ref = nullptr;
    uint32_t nref = 0;
    const sdbfx_set* sample = nullptr;   // the reference set's first digests, for autotuning
    sdbfx_set* own_sample = nullptr;     // owned copy when the sample is not the whole set
    uint64_t batch_bytes = 0;
    uint32_t batch_max = 0;
    bool fast = false;         // the reference set is folded (sdbfx_set_fold)
    sdbfx_ctx* cd = nullptr;   // digest context
    sdbfx_ctx* cc = nullptr;   // compare context

    std::mutex mu;
    std::condition_variable cv_queue, cv_handoff, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::deque<Batch> handoff;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::string err;
    uint64_t nbatches = 0, nqueries = 0, nmfma = 0, ndeferred = 0, nrefused = 0;
    int8_t tuned[64];   // per size class: 0 untuned, SDBFX_CMP_VALU or SDBFX_CMP_MFMA
    std::thread td, tc;

    void set_error(const std::string& e)
    {
        std::lock_guard<std::mutex> g(mu);
        if (err.empty())
            err = e;
    }

    void digest_loop()
    {
        for (;;) {
            Batch b;
            {
                std::unique_lock<std::mutex> g(mu);
                // a new batch only when the compare stage has room for it
                cv_queue.wait(g, [&] { return stop || (!queue.empty() && handoff.size() < 1); });
                if (stop)
                    return;
                uint64_t bytes = 0;
                while (!queue.empty() && b.reqs.size() < batch_max) {
                    const auto& r = queue.front();
                    if (!b.reqs.empty() && bytes + r->data.size() > batch_bytes)
                        break;
                    bytes += r->data.size();
                    b.reqs.push_back(r);
                    queue.pop_front();
                }
            }
            const uint32_t n = (uint32_t)b.reqs.size();
            std::vector<const uint8_t*> bufs(n);
            std::vector<uint64_t> lens(n);
            std::vector<const char*> names(n);
            std::vector<uint32_t> dd(n);
            for (uint32_t i = 0; i < n; i++) {
                bufs[i] = b.reqs[i]->data.data();
                lens[i] = b.reqs[i]->data.size();
                names[i] = b.reqs[i]->name.c_str();
                dd[i] = b.reqs[i]->dd;
            }
            b.status.assign(n, 0);
            int rc = sdbfx_set_create(cd, &b.set);
            uint32_t m = n;   // the queries digested: [0, m) of the batch
            bool retried = false;
            for (;;) {
                if (rc == 0)
                    rc = sdbfx_set_digest(cd, b.set, m, bufs.data(), lens.data(), names.data(), dd.data(),
                                          b.status.data());
                if (rc != SDBFX_EBUDGET)
                    break;
                // the scratch that earlier batches (or this failed attempt)
                // left may be what is in the way: free it, then try the same
                // queries once more before splitting them
                if ((rc = sdbfx_ctx_trim(cd)) != 0)
                    break;
                if (retried && m == 1) {
                    rc = SDBFX_EBUDGET;
                    break;
                }
                if (retried)
                    m = (m + 1) / 2;
                retried = true;
            }
            if (m < n) {
                // deferred: back to the front of the queue, in order
                std::lock_guard<std::mutex> g(mu);
                for (uint32_t i = n; i-- > m;)
                    queue.push_front(b.reqs[i]);
                ndeferred += n - m;
                b.reqs.resize(m);
                b.status.resize(m);
            }
            // a folded reference set (fast mode) is compared with folded queries
            if (rc == 0 && fast)
                rc = sdbfx_set_fold(cd, b.set);
            if (rc == SDBFX_EBUDGET) {
                b.status[0] = SDBFX_EBUDGET;   // a single query that does not fit
                std::lock_guard<std::mutex> g(mu);
                nrefused++;
            } else if (rc != 0) {
                set_error(std::string("digest: ") + sdbfx_last_error(cd));
                for (uint32_t i = 0; i < m; i++)
                    b.status[i] = -1;
            }
            for (auto& r : b.reqs)
                std::vector<uint8_t>().swap(r->data);   // input no longer needed
            if (m < n)
                cv_queue.notify_one();
            std::lock_guard<std::mutex> g(mu);
            handoff.push_back(std::move(b));
            cv_handoff.notify_one();
        }
    }

    // True when the MFMA kernel's padded work on (batch x sample) is under
    // half the VALU kernel's, by the engine's cost model (sdbfx_host.cpp,
    // use_mfma: 2 vs 8 SIMD cycles per padded pair slot; query filters
    // padded to 128 and 64 rows, target filters to 128 and 1); then the
    // trial is skipped.  The batch's digests all face every sample digest,
    // so the instances per target digest are the batch's filter count.
    bool clear_mfma(const sdbfx_set* q) const
    {
        const double nq = (double)sdbfx_set_filter_count(q);
        double valu = 0, mfma = 0;
        sdbfx_digest_info info;
        for (uint32_t i = 0; i < sdbfx_set_size(sample); i++) {
            if (sdbfx_set_info(sample, i, &info) != 0)
                return false;
            const double t = info.bf_count;
            valu += 8.0 * std::ceil(nq / 64) * 64 * t;
            mfma += 2.0 * std::ceil(nq / 128) * 128 * std::ceil(t / 128) * 128;
        }
        return mfma < 0.5 * valu;
    }

    // q's digests [q0, q1) against t, split into ranges that fit the compare
    // context's budget; `refused` receives the digests that do not fit alone
    int compare_fit(const sdbfx_set* q, uint32_t q0, uint32_t q1, const sdbfx_set* t, int32_t* out,
                    std::vector<uint32_t>& refused)
    {
        if (q0 >= q1)
            return 0;
        const uint32_t nt = sdbfx_set_size(t);
        int rc = sdbfx_compare_to_range(cc, q, q0, q1 - q0, t, out + (size_t)q0 * nt);
        if (rc == SDBFX_EBUDGET) {
            rc = sdbfx_ctx_trim(cc);
            if (rc == 0)
                rc = sdbfx_compare_to_range(cc, q, q0, q1 - q0, t, out + (size_t)q0 * nt);
        }
        if (rc != SDBFX_EBUDGET)
            return rc;
        if (q1 - q0 == 1) {
            refused.push_back(q0);
            return 0;
        }
        const uint32_t mid = q0 + (q1 - q0) / 2;
        if ((rc = compare_fit(q, q0, mid, t, out, refused)))
            return rc;
        return compare_fit(q, mid, q1, t, out, refused);
    }

    int timed_compare(const sdbfx_set* q, const sdbfx_set* t, int method, int32_t* out, double* seconds,
                      std::vector<uint32_t>* refused = nullptr)
    {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = sdbfx_set_compare_method(cc, method);
        std::vector<uint32_t> none;
        if (rc == 0)
            rc = compare_fit(q, 0, sdbfx_set_size(q), t, out, refused ? *refused : none);
        if (!refused && !none.empty())
            rc = SDBFX_EBUDGET;
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    }

    void compare_loop()
    {
        for (;;) {
            Batch b;
            {
                std::unique_lock<std::mutex> g(mu);
                cv_handoff.wait(g, [&] { return stop || !handoff.empty(); });
                if (stop && handoff.empty())
                    return;
                b = std::move(handoff.front());
                handoff.pop_front();
                cv_queue.notify_one();   // the digest stage may start the next batch
            }
            const uint32_t n = (uint32_t)b.reqs.size();
            const uint32_t nd = b.set ? sdbfx_set_size(b.set) : 0;
            std::vector<int32_t> sc((size_t)nd * nref);
            std::vector<uint32_t> refused;   // digests whose comparison does not fit the budget
            int rc = 0, method = SDBFX_CMP_AUTO;
            if (nd && nref) {
                const uint64_t nf = sdbfx_set_filter_count(b.set);
                const int cls = nf ? 63 - __builtin_clzll(nf) : 0;
                method = tuned[cls];
                if (method == 0 && clear_mfma(b.set)) {
                    method = SDBFX_CMP_MFMA;
                    tuned[cls] = (int8_t)method;
                }
                if (method == 0) {
                    // time both kernels on this batch against the sample;
                    // keep the faster
                    const size_t ns = (size_t)nd * sdbfx_set_size(sample);
                    std::vector<int32_t> sv(ns), sm(ns);
                    double tv = 0, tm = 0;
                    rc = timed_compare(b.set, sample, SDBFX_CMP_VALU, sv.data(), &tv);
                    if (rc == 0)
                        rc = timed_compare(b.set, sample, SDBFX_CMP_MFMA, sm.data(), &tm);
                    if (rc == 0 && sv != sm)
                        set_error("compare: the two kernels disagree");
                    if (rc == SDBFX_EBUDGET) {
                        rc = 0;   // the trial does not fit: no tuning for this batch
                        method = SDBFX_CMP_AUTO;
                    } else {
                        method = tm <= tv ? SDBFX_CMP_MFMA : SDBFX_CMP_VALU;
                        tuned[cls] = (int8_t)method;
                    }
                }
                if (rc == 0) {
                    double t;
                    rc = timed_compare(b.set, ref, method, sc.data(), &t, &refused);
                }
            }
            if (rc != 0)
                set_error(std::string("compare: ") + sdbfx_last_error(cc));
            const uint64_t t = now_ns();
            {
                std::lock_guard<std::mutex> g(mu);
                uint32_t k = 0;   // row of the next digested request
                size_t rf = 0;    // next refused row
                for (uint32_t i = 0; i < n; i++) {
                    Request& r = *b.reqs[i];
                    r.status = rc != 0 ? -1 : b.status[i];
                    if (b.status[i] == 0) {
                        if (rf < refused.size() && refused[rf] == k) {
                            r.status = SDBFX_EBUDGET;
                            rf++;
                            nrefused++;
                        } else if (rc == 0) {
                            r.scores.assign(sc.begin() + (size_t)k * nref, sc.begin() + (size_t)(k + 1) * nref);
                        }
                        k++;
                    }
                    r.t_done = t;
                    r.done = true;
                }
                nbatches++;
                nqueries += n;
                nmfma += method == SDBFX_CMP_MFMA ? 1 : 0;   // AUTO counts as not tuned
            }
            cv_done.notify_all();
            if (b.set)
                sdbfx_set_destroy(b.set);
        }
    }
};

extern "C" {

namespace {

// The autotuning sample: the reference set's first whole digests up to
// SERVE_TUNE_SAMPLE_FILTERS filters, copied into a set of the compare
// context (or the reference set itself when it is not much larger).
int make_sample(sdbfx_server* s)
{
    uint32_t k = 0;
    uint64_t nf = 0;
    sdbfx_digest_info info;
    while (k < s->nref && nf < SERVE_TUNE_SAMPLE_FILTERS) {
        if (sdbfx_set_info(s->ref, k, &info) != 0)
            return -1;
        nf += info.bf_count;
        k++;
    }
    if (k == s->nref || sdbfx_set_filter_count(s->ref) <= 2 * nf) {
        s->sample = s->ref;
        return 0;
    }
    if (sdbfx_set_create(s->cc, &s->own_sample) != 0)
        return -1;
    std::vector<uint8_t> filt;
    std::vector<uint16_t> ec;
    for (uint32_t i = 0; i < k; i++) {
        if (sdbfx_set_info(s->ref, i, &info) != 0)
            return -1;
        filt.resize((size_t)info.bf_count * 256);
        ec.resize(info.bf_count);
        if (sdbfx_set_read(s->cc, s->ref, i, filt.data(), ec.data(), nullptr) != 0 ||
            sdbfx_set_add(s->cc, s->own_sample, info.name, info.orig_size, info.bf_count, info.last_count,
                          info.dd_block_size, info.max_elem, filt.data(), ec.data()) != 0)
            return -1;
    }
    s->sample = s->own_sample;
    return 0;
}

}  // namespace

int sdbfx_serve_create(int device, const sdbfx_set* ref, uint64_t batch_bytes, uint32_t batch_max,
                       uint64_t mem_budget, sdbfx_server** out)
{
    if (!out || !ref || batch_bytes == 0 || batch_max == 0)
        return -1;
    *out = nullptr;
    auto* s = new sdbfx_server();
    memset(s->tuned, 0, sizeof(s->tuned));
    s->ref = ref;
    s->nref = sdbfx_set_size(ref);
    s->fast = sdbfx_set_folded(ref) == 1;
    s->batch_bytes = batch_bytes;
    s->batch_max = batch_max;
    if (sdbfx_ctx_create(device, mem_budget / 2, &s->cd) != 0 || sdbfx_ctx_create(device, mem_budget / 2, &s->cc) != 0) {
        if (s->cd)
            sdbfx_ctx_destroy(s->cd);
        delete s;
        return -1;
    }
    // the sample copies folded filters of a folded reference: fold it too
    // (folding is idempotent) so that it is compared in the same mode
    if (make_sample(s) != 0 || (s->fast && s->own_sample && sdbfx_set_fold(s->cc, s->own_sample) != 0)) {
        if (s->own_sample)
            sdbfx_set_destroy(s->own_sample);
        sdbfx_ctx_destroy(s->cd);
        sdbfx_ctx_destroy(s->cc);
        delete s;
        return -1;
    }
    s->td = std::thread([s] { s->digest_loop(); });
    s->tc = std::thread([s] { s->compare_loop(); });
    *out = s;
    return 0;
}

void sdbfx_serve_destroy(sdbfx_server* s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_queue.notify_all();
    s->cv_handoff.notify_all();
    s->td.join();
    s->tc.join();
    for (auto& b : s->handoff)
        if (b.set)
            sdbfx_set_destroy(b.set);
    if (s->own_sample)
        sdbfx_set_destroy(s->own_sample);
    sdbfx_ctx_destroy(s->cd);
    sdbfx_ctx_destroy(s->cc);
    delete s;
}

const char* sdbfx_serve_error(sdbfx_server* s)
{
    if (!s)
        return "null server";
    std::lock_guard<std::mutex> g(s->mu);
    return s->err.c_str();
}

uint32_t sdbfx_serve_ref_size(const sdbfx_server* s) { return s ? s->nref : 0; }

int sdbfx_serve_submit(sdbfx_server* s, const uint8_t* buf, uint64_t len, const char* name, uint32_t dd_block,
                       uint64_t* ticket)
{
    if (!s || (!buf && len) || !name || !ticket)
        return -1;
    auto r = std::make_shared<Request>();
    r->data.assign(buf, buf + len);
    r->name = name;
    r->dd = dd_block;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop)
            return -1;
        r->ticket = s->next_ticket++;
        r->t_submit = now_ns();
        s->live[r->ticket] = r;
        s->queue.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_queue.notify_one();
    return 0;
}

// Shared by poll and wait: hands the result over and forgets the ticket.
static int take(sdbfx_server* s, std::unordered_map<uint64_t, std::shared_ptr<Request>>::iterator it,
                int32_t* scores, int32_t* status, uint64_t* latency_ns)
{
    Request& r = *it->second;
    if (status)
        *status = r.status;
    if (latency_ns)
        *latency_ns = r.t_done - r.t_submit;
    if (scores && r.status == 0)
        memcpy(scores, r.scores.data(), r.scores.size() * sizeof(int32_t));
    s->live.erase(it);
    return 0;
}

int sdbfx_serve_poll(sdbfx_server* s, uint64_t ticket, int32_t* scores, int32_t* status, uint64_t* latency_ns,
                     int* done)
{
    if (!s || !done)
        return -1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end())
        return -1;
    *done = it->second->done ? 1 : 0;
    return *done ? take(s, it, scores, status, latency_ns) : 0;
}

int sdbfx_serve_wait(sdbfx_server* s, uint64_t ticket, int32_t* scores, int32_t* status, uint64_t* latency_ns)
{
    if (!s)
        return -1;
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end())
        return -1;
    std::shared_ptr<Request> r = it->second;
    s->cv_done.wait(g, [&] { return r->done; });
    return take(s, s->live.find(ticket), scores, status, latency_ns);
}

int sdbfx_serve_admission(sdbfx_server* s, uint64_t* deferred, uint64_t* refused)
{
    if (!s)
        return -1;
    std::lock_guard<std::mutex> g(s->mu);
    if (deferred)
        *deferred = s->ndeferred;
    if (refused)
        *refused = s->nrefused;
    return 0;
}

int sdbfx_serve_stats(sdbfx_server* s, uint64_t* batches, uint64_t* queries, uint64_t* mfma_batches)
{
    if (!s)
        return -1;
    std::lock_guard<std::mutex> g(s->mu);
    if (batches)
        *batches = s->nbatches;
    if (queries)
        *queries = s->nqueries;
    if (mfma_batches)
        *mfma_batches = s->nmfma;
    return 0;
}

}  // extern "C"
