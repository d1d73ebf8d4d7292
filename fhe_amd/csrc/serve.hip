// A serving loop over resident keys and ciphertexts: jobs arrive over time
// (the caller gives each an arrival time), wait in queues per (engine, job
// kind, level), and are launched in batches as they arrive - continuous
// batching: a batch is formed from whatever has arrived whenever an engine
// has room for another batch on the device, so requests that arrive while
// others run join the next batch.  Each engine (a context: ring degree,
// prime chain, keys) has its own stream, so batches of different ring
// degrees overlap on the device.  A batch is admitted only when the free
// HBM (hipMemGetInfo) minus its estimated working set stays above a
// reserve; otherwise it waits for running batches to finish.
//
// The jobs are compositions of the engine operations the reference defines
// (CKKSEngine::multiply(ct, ct, rlk) and sum_1d_by_row,
// src/ckks/engine_multiply.cpp:134-147, src/ckks/engine_advanced.cpp:
// 549-577; evaluate_polynomial, :365-387), each job's limbs equal to the
// same composition run alone: a batch runs its jobs side by side through
// the batched device operations, which compute each ciphertext exactly as
// alone.  The reference has no serving loop.
#include <chrono>
#include <deque>
#include <thread>

#include "text_engine.h"

namespace fhe_host {
namespace {

using Clock = std::chrono::steady_clock;

struct Batch {
  int engine = 0;
  std::vector<int> jobs;
  hipEvent_t done = nullptr;
};

// Device bytes a batch of `count` jobs at `limbs` limbs may take beyond its
// inputs: the key switch's extended limbs, accumulators and trio for each
// ciphertext, and the texts of the composition (power basis or the
// rotation chain), in limbs of N words.
size_t batch_bytes(const fhe_context *ctx, int kind, int count, int limbs) {
  const auto &P = ctx->params;
  const int ext = limbs + P.alpha;
  const int parts = (limbs + P.alpha - 1) / P.alpha;
  const int64_t ks = int64_t(parts) * ext + 2 * ext + 3 * limbs;
  const int64_t texts = kind == FHE_JOB_POLY ? 9 * 2 * limbs : 6 * 2 * limbs;
  return size_t(count) * size_t(ks + texts) * size_t(P.n) * 8;
}

} // namespace
} // namespace fhe_host

extern "C" int fhe_serve(const fhe_serve_engine *engines, int engine_count,
                         fhe_job *jobs, int job_count,
                         const fhe_serve_options *options,
                         fhe_serve_stats *stats) {
  return guarded([&] {
    require(engines && jobs && options && stats, "null argument");
    require(engine_count > 0 && job_count >= 0, "Serve: empty");
    require(options->max_batch > 0 && options->max_in_flight > 0,
            "Serve: max_batch and max_in_flight must be positive");
    std::vector<fhe_bootstrap_keys> keys(engine_count);
    for (int g = 0; g < engine_count; g++) {
      const auto &E = engines[g];
      require(E.ctx && E.rlk, "null argument");
      check_rotation_key(&E.rotation);
      keys[g] = rotation_keys_of(E.rotation);
    }
    for (int j = 0; j < job_count; j++) {
      const fhe_job &J = jobs[j];
      require(J.engine >= 0 && J.engine < engine_count,
              "Serve: job engine out of range");
      require(J.kind == FHE_JOB_DOT || J.kind == FHE_JOB_POLY,
              "Serve: unknown job kind");
      require(J.a && J.out && (J.kind != FHE_JOB_DOT || J.b),
              "null argument");
      const auto *ctx = engines[J.engine].ctx;
      require(J.level >= 1 && J.level <= ctx->params.full_level,
              "Serve: job level out of range");
    }
    *stats = fhe_serve_stats{};
    stats->min_free_bytes = ~size_t(0);

    std::vector<int> order(job_count);
    for (int j = 0; j < job_count; j++)
      order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return jobs[x].arrival_ms < jobs[y].arrival_ms;
    });
    // waiting jobs per (engine, kind, level), in arrival order
    std::map<std::tuple<int, int, int>, std::deque<int>> waiting;
    std::vector<fhe_host::Batch> running;
    std::vector<int> in_flight(engine_count, 0);
    const auto t0 = fhe_host::Clock::now();
    auto now_ms = [&] {
      return std::chrono::duration<double, std::milli>(fhe_host::Clock::now() -
                                                       t0)
          .count();
    };
    int next = 0, done = 0;
    try {
      while (done < job_count) {
        bool progress = false;
        // arrivals
        const double now = now_ms();
        while (next < job_count && jobs[order[next]].arrival_ms <= now) {
          const fhe_job &J = jobs[order[next]];
          waiting[{J.engine, J.kind, J.level}].push_back(order[next]);
          next++;
          progress = true;
        }
        // completions
        for (size_t b = 0; b < running.size();) {
          const hipError_t q = hipEventQuery(running[b].done);
          if (q == hipErrorNotReady) {
            b++;
            continue;
          }
          hip_check(q, "Serve: batch");
          const double t = now_ms();
          for (int j : running[b].jobs)
            jobs[j].done_ms = t;
          done += int(running[b].jobs.size());
          in_flight[running[b].engine]--;
          (void)hipEventDestroy(running[b].done);
          running.erase(running.begin() + long(b));
          progress = true;
        }
        // dispatch: every engine with room takes the queue whose first job
        // arrived first
        for (int g = 0; g < engine_count; g++) {
          if (in_flight[g] >= options->max_in_flight)
            continue;
          auto best = waiting.end();
          for (auto it = waiting.begin(); it != waiting.end(); ++it)
            if (std::get<0>(it->first) == g && !it->second.empty() &&
                (best == waiting.end() ||
                 jobs[it->second.front()].arrival_ms <
                     jobs[best->second.front()].arrival_ms))
              best = it;
          if (best == waiting.end())
            continue;
          const int kind = std::get<1>(best->first);
          const int level = std::get<2>(best->first);
          const auto &E = engines[g];
          fhe_context *ctx = E.ctx;
          const int limbs = ctx->params.limbs_at(ctx->moduli_id_of(level, false));
          const int count = std::min<int>(options->max_batch,
                                          int(best->second.size()));
          size_t free_b = 0, total_b = 0;
          hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
          stats->min_free_bytes = std::min(stats->min_free_bytes, free_b);
          const size_t need = fhe_host::batch_bytes(ctx, kind, count, limbs);
          if (free_b < need + options->reserve_bytes) {
            // wait for memory: only when something can free it
            stats->deferred_admissions++;
            if (!running.empty())
              continue;
          }
          fhe_host::Batch batch;
          batch.engine = g;
          for (int m = 0; m < count; m++) {
            batch.jobs.push_back(best->second.front());
            best->second.pop_front();
          }
          const double admitted = now_ms();
          fhe_host::TextEngine e(ctx, E.rlk, E.rlk_moduli_id);
          e.sync_on_exit = false;
          e.count = count;
          e.keys = &keys[g];
          const int64_t words = int64_t(2) * limbs * ctx->n();
          fhe_host::Text a = e.alloc(2, level, 0, count);
          fhe_host::Text b;
          for (int m = 0; m < count; m++)
            hip_check(hipMemcpyAsync(a.data + m * words,
                                     jobs[batch.jobs[m]].a, size_t(words) * 8,
                                     hipMemcpyDeviceToDevice, ctx->stream),
                      "Serve: gather");
          if (kind == FHE_JOB_DOT) {
            b = e.alloc(2, level, 0, count);
            for (int m = 0; m < count; m++)
              hip_check(hipMemcpyAsync(b.data + m * words,
                                       jobs[batch.jobs[m]].b,
                                       size_t(words) * 8,
                                       hipMemcpyDeviceToDevice, ctx->stream),
                        "Serve: gather");
          }
          const fhe_host::Text r =
              kind == FHE_JOB_DOT
                  ? fhe_host::sum_1d_by_row(e, e.multiply(a, b, true))
                  : fhe_host::evaluate_polynomial(
                        e, a,
                        std::vector<double>(E.coefficients,
                                            E.coefficients +
                                                E.coefficient_count));
          const int64_t rw = e.words(r) / count;
          for (int m = 0; m < count; m++) {
            fhe_job &J = jobs[batch.jobs[m]];
            hip_check(hipMemcpyAsync(J.out, r.data + m * rw, size_t(rw) * 8,
                                     hipMemcpyDeviceToDevice, ctx->stream),
                      "Serve: scatter");
            J.out_level = r.level;
            J.out_flags = r.flags;
            J.admitted_ms = admitted;
          }
          hip_check(hipEventCreateWithFlags(&batch.done,
                                            hipEventDisableTiming),
                    "hipEventCreate");
          hip_check(hipEventRecord(batch.done, ctx->stream), "hipEventRecord");
          ctx->check_launch("Serve");
          running.push_back(std::move(batch));
          in_flight[g]++;
          stats->batches++;
          stats->max_batch_jobs = std::max(stats->max_batch_jobs, count);
          progress = true;
        }
        if (!progress) {
          // nothing arrived, finished or launched: sleep until the next
          // arrival when nothing runs, else poll again shortly
          double wait_ms = running.empty() && next < job_count
                               ? jobs[order[next]].arrival_ms - now_ms()
                               : 0.02;
          wait_ms = std::min(std::max(wait_ms, 0.0), 1.0);
          std::this_thread::sleep_for(
              std::chrono::duration<double, std::milli>(wait_ms));
        }
      }
    } catch (...) {
      for (auto &b : running) {
        (void)hipEventSynchronize(b.done);
        (void)hipEventDestroy(b.done);
      }
      throw;
    }
    stats->wall_ms = now_ms();
  });
}
