// Batched and served commit+open jobs on one GPU.
//
// A server owns `workers` host threads, each with its own bitz_ctx (device
// memory pool and HIP stream), so the kernels, transcript work and copies of
// different jobs overlap on the device. Jobs are admitted in submission order
// when a worker is free and the device-memory footprint of the jobs in flight
// plus the next job fits the budget (a job larger than the whole budget runs
// alone). Each job is the reference's commit + standalone opening
// (Pcs::commit, src/bitz/pcs.rs:270-284; standalone_evaluation /
// prove_standalone, src/piop/spartan/protocol/bitz_opener.rs:643-707): the
// same calls, on a context of its own, as one job run by itself, hence the
// same proof bytes.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bitz_hip.h"

namespace bitz {
void set_error(const std::string& msg);  // ctx.hip
}

namespace {

using Clock = std::chrono::steady_clock;

struct Worker {
    bitz_ctx* ctx = nullptr;
    std::thread th;
};

}  // namespace

struct bitz_server {
    int device = 0;
    uint64_t budget = 0;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<bitz_job*> queue;
    uint64_t in_flight_bytes = 0;
    uint32_t in_flight = 0;
    uint64_t pending = 0;  // submitted and not finished
    bool stop = false;
    std::vector<Worker> workers;
};

extern "C" uint64_t bitz_job_footprint(uint32_t t, uint32_t s, uint32_t log_batch, uint32_t log_inv_rate) {
    const int m = (int)(t + s), mp = m - 7;
    if (t < 7 || mp < 1 || log_batch > 6 || log_inv_rate > 8) return 0;
    double b = 16.0 * std::ldexp(1.0, m);                                   // product-tree storage
    b += 16.0 * std::ldexp(1.0, mp + (int)log_inv_rate);                    // L0 codeword
    b += 64.0 * std::ldexp(1.0, mp - (int)log_batch + (int)log_inv_rate);   // L0 Merkle tree
    b += 6.0 * 16.0 * std::ldexp(1.0, mp);                                  // rows, packing, Ligerito tables
    b += 16.0 * (std::ldexp(1.0, (int)t + 1) + std::ldexp(1.0, (int)t + 6));  // bit-pattern tables
    b += 64.0 * 1048576.0;                                                  // eq tables, scratch
    return (uint64_t)b;
}

namespace {

double ms_since(const bitz_server* srv) {
    return std::chrono::duration<double, std::milli>(Clock::now() - srv->t0).count();
}

// the committed rows of a u64 multiplication job, packed on the device
int mul_rows(bitz_ctx* ctx, const bitz_job* j, void** rows) {
    void *xd = nullptr, *yd = nullptr;
    const uint64_t bytes = 8 * j->mul_n, words = 1ull << (j->t + j->s - 6);
    int rc = bitz_dev_alloc(ctx, bytes, &xd);
    if (!rc) rc = bitz_dev_alloc(ctx, bytes, &yd);
    if (!rc) rc = bitz_dev_alloc(ctx, 8 * words, rows);
    if (!rc) rc = bitz_dev_upload(ctx, xd, j->mul_x, bytes);
    if (!rc) rc = bitz_dev_upload(ctx, yd, j->mul_y, bytes);
    if (!rc) rc = bitz_mul_bit_rows_dev(ctx, 64, j->mul_n, 0, xd, yd, nullptr, nullptr, *rows, words);
    if (xd) bitz_dev_free(ctx, xd);  // stream-ordered: the packing has read them
    if (yd) bitz_dev_free(ctx, yd);
    return rc;
}

// the u64 multiplication proof of one job (bitz_prove_mul_u64), its outer
// transcript state in claimed[0..16) (the first half of the digest)
int run_mul_proof_job(bitz_ctx* ctx, bitz_job* j) {
    void *xd = nullptr, *yd = nullptr;
    const uint64_t bytes = 8 * j->mul_n;
    uint8_t state[32];
    int rc = bitz_dev_alloc(ctx, bytes, &xd);
    if (!rc) rc = bitz_dev_alloc(ctx, bytes, &yd);
    if (!rc) rc = bitz_dev_upload(ctx, xd, j->mul_x, bytes);
    if (!rc) rc = bitz_dev_upload(ctx, yd, j->mul_y, bytes);
    if (!rc)
        rc = bitz_prove_mul_u64(ctx, xd, yd, nullptr, nullptr, j->mul_n, 0, j->root, state, j->proof,
                                j->proof_capacity, &j->proof_len);
    if (xd) bitz_dev_free(ctx, xd);
    if (yd) bitz_dev_free(ctx, yd);
    if (rc) return rc;
    memcpy(j->claimed, state, 16);
    if (!j->verify) return rc;
    rc = bitz_verify_mul_u64(ctx, j->root, j->mul_n, 0, j->proof, j->proof_len, nullptr);
    j->verified = rc == BITZ_OK ? 1 : 0;
    return rc;
}

// the SHA-256 chain proof of one job (bitz_prove_sha256_chain), its outer
// transcript state in claimed[0..16)
int run_chain_job(bitz_ctx* ctx, bitz_job* j) {
    uint8_t state[32];
    int rc = bitz_prove_sha256_chain(ctx, j->chain_blocks, j->chain_log_n, nullptr, j->root, j->chain_digest, state,
                                     j->proof, j->proof_capacity, &j->proof_len);
    if (rc) return rc;
    memcpy(j->claimed, state, 16);
    if (!j->verify) return rc;
    rc = bitz_verify_sha256_chain(ctx, j->chain_blocks, j->chain_log_n, nullptr, j->chain_digest, j->root, j->proof,
                                  j->proof_len, nullptr);
    j->verified = rc == BITZ_OK ? 1 : 0;
    return rc;
}

// commit + (evaluation) + proof (+ verification) of one job on one context
int run_job(bitz_ctx* ctx, bitz_job* j) {
    if (j->kind == BITZ_JOB_MUL_U64_PROOF) return run_mul_proof_job(ctx, j);
    if (j->kind == BITZ_JOB_SHA256_CHAIN) return run_chain_job(ctx, j);
    bitz_commit_handle* h = nullptr;
    int rc;
    if (j->kind == BITZ_JOB_MUL_U64) {
        void* rows = nullptr;
        rc = mul_rows(ctx, j, &rows);
        if (!rc) rc = bitz_commit_device_rows(ctx, rows, j->t, j->s, j->log_batch, j->log_inv_rate, j->root, &h);
        if (rows) bitz_dev_free(ctx, rows);  // the handle holds its own copy
    } else {
        rc = j->rows_on_device ? bitz_commit_device_rows(ctx, j->rows, j->t, j->s, j->log_batch, j->log_inv_rate,
                                                         j->root, &h)
                               : bitz_commit(ctx, (const uint64_t*)j->rows, j->t, j->s, j->log_batch,
                                             j->log_inv_rate, j->root, &h);
    }
    if (rc) return rc;
    if (!j->have_claim) {
        rc = bitz_standalone_evaluation(ctx, h, j->claimed);
        if (rc) {
            bitz_commit_free(h);
            return rc;
        }
    }
    rc = bitz_prove_standalone(ctx, h, j->claimed, j->proof, j->proof_capacity, &j->proof_len);
    bitz_commit_free(h);
    if (rc || !j->verify) return rc;
    rc = bitz_verify_standalone(ctx, j->root, j->t, j->s, j->log_inv_rate, j->log_batch, j->claimed, j->proof,
                                j->proof_len);
    j->verified = rc == BITZ_OK ? 1 : 0;
    return rc;
}

void worker_loop(bitz_server* srv, Worker* w) {
    for (;;) {
        bitz_job* j = nullptr;
        uint64_t fp = 0;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] {
                if (srv->stop) return true;
                if (srv->queue.empty()) return false;
                bitz_job* head = srv->queue.front();
                uint64_t need = bitz_job_footprint(head->t, head->s, head->log_batch, head->log_inv_rate);
                return srv->in_flight == 0 || srv->in_flight_bytes + need <= srv->budget;
            });
            if (srv->stop && srv->queue.empty()) return;
            if (srv->queue.empty()) continue;
            j = srv->queue.front();
            srv->queue.pop_front();
            fp = bitz_job_footprint(j->t, j->s, j->log_batch, j->log_inv_rate);
            srv->in_flight_bytes += fp;
            srv->in_flight++;
        }
        j->start_ms = ms_since(srv);
        j->status = run_job(w->ctx, j);
        if (j->status) {
            const char* e = bitz_last_error();
            size_t n = std::min(sizeof(j->error) - 1, strlen(e));
            memcpy(j->error, e, n);
            j->error[n] = 0;
        }
        j->end_ms = ms_since(srv);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->in_flight_bytes -= fp;
            srv->in_flight--;
            srv->pending--;
            j->done = 1;
        }
        srv->cv_work.notify_all();
        srv->cv_done.notify_all();
    }
}

}  // namespace

extern "C" int bitz_server_create(int device, uint32_t workers, uint64_t hbm_budget, bitz_server** out) {
    if (!out || workers == 0 || workers > 64) {
        bitz::set_error("bitz_server_create: need out and 1 <= workers <= 64");
        return BITZ_ERR_ARG;
    }
    auto* srv = new bitz_server();
    srv->device = device;
    srv->budget = hbm_budget;
    srv->t0 = Clock::now();
    srv->workers.resize(workers);
    for (auto& w : srv->workers) {
        int rc = bitz_ctx_create(device, &w.ctx);
        if (rc) {
            for (auto& v : srv->workers)
                if (v.ctx) bitz_ctx_destroy(v.ctx);
            delete srv;
            return rc;
        }
    }
    for (auto& w : srv->workers) w.th = std::thread(worker_loop, srv, &w);
    *out = srv;
    return BITZ_OK;
}

extern "C" int bitz_server_submit(bitz_server* srv, bitz_job* job) {
    if (!srv || !job || !job->proof) {
        bitz::set_error("bitz_server_submit: null argument");
        return BITZ_ERR_ARG;
    }
    if (job->kind == BITZ_JOB_MUL_U64 || job->kind == BITZ_JOB_MUL_U64_PROOF) {
        uint32_t lay[4];
        if (!job->mul_x || !job->mul_y) {
            bitz::set_error("bitz_server_submit: null multiplication operands");
            return BITZ_ERR_ARG;
        }
        if (int rc = bitz_mul_layout(64, job->mul_n, 0, lay)) return rc;
        job->t = lay[2];
        job->s = lay[1];
    } else if (job->kind == BITZ_JOB_SHA256_CHAIN) {
        if (!job->chain_blocks || job->chain_log_n < 7 || job->chain_log_n > 16) {
            bitz::set_error("bitz_server_submit: chain jobs need 2^7..2^16 host blocks");
            return BITZ_ERR_ARG;
        }
        // the product tensor's grid (the largest buffers the job holds)
        job->t = std::min(job->chain_log_n, 13u);
        job->s = 15 + job->chain_log_n - job->t;
        job->log_batch = 4;
        job->log_inv_rate = 1;
    } else if (job->kind != BITZ_JOB_ROWS || !job->rows) {
        bitz::set_error("bitz_server_submit: null rows or unknown job kind");
        return BITZ_ERR_ARG;
    }
    job->verified = 0;
    if (bitz_job_footprint(job->t, job->s, job->log_batch, job->log_inv_rate) == 0) {
        bitz::set_error("bitz_server_submit: job shape rejected");
        return BITZ_ERR_SHAPE;
    }
    job->done = 0;
    job->status = 0;
    job->error[0] = 0;
    job->proof_len = 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) {
            bitz::set_error("bitz_server_submit: server is stopping");
            return BITZ_ERR_ARG;
        }
        job->submit_ms = ms_since(srv);
        srv->queue.push_back(job);
        srv->pending++;
    }
    srv->cv_work.notify_all();
    return BITZ_OK;
}

extern "C" int bitz_server_done(bitz_server* srv, const bitz_job* job) {
    if (!srv || !job) return BITZ_ERR_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    return job->done ? 1 : 0;
}

extern "C" int bitz_server_wait_all(bitz_server* srv) {
    if (!srv) return BITZ_ERR_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->pending == 0; });
    return BITZ_OK;
}

extern "C" double bitz_server_clock_ms(bitz_server* srv) { return srv ? ms_since(srv) : 0.0; }

extern "C" void bitz_server_destroy(bitz_server* srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& w : srv->workers)
        if (w.th.joinable()) w.th.join();
    for (auto& w : srv->workers)
        if (w.ctx) bitz_ctx_destroy(w.ctx);
    delete srv;
}

extern "C" int bitz_batch_run(int device, uint32_t workers, uint64_t hbm_budget, bitz_job* jobs, uint32_t n) {
    if (n && !jobs) return BITZ_ERR_ARG;
    bitz_server* srv = nullptr;
    int rc = bitz_server_create(device, workers, hbm_budget, &srv);
    if (rc) return rc;
    // largest jobs first: the tail of the batch is short jobs that fill gaps
    std::vector<bitz_job*> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = &jobs[i];
    std::stable_sort(order.begin(), order.end(), [](const bitz_job* a, const bitz_job* b) {
        return bitz_job_footprint(a->t, a->s, a->log_batch, a->log_inv_rate) >
               bitz_job_footprint(b->t, b->s, b->log_batch, b->log_inv_rate);
    });
    for (auto* j : order)
        if ((rc = bitz_server_submit(srv, j))) break;
    bitz_server_wait_all(srv);
    bitz_server_destroy(srv);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; i++)
        if (jobs[i].status) return jobs[i].status;
    return BITZ_OK;
}
