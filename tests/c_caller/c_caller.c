/*
 * A plain C caller of the tsgpu C ABI (include/tsgpu.h), the way a C
 * program or a foreign-function binding would use it: context, job
 * (Session::builder().build()), run with a progress callback
 * (Session::run), download (GeneratedImage), destroy; plus one validation
 * error (Error::InvalidRange) and the many-job batch API.
 *
 *   c_caller EXAMPLE.rgba W H OUT_SIDE SEED OUT_PREFIX
 *
 * reads a W x H RGBA8 example, writes OUT_PREFIX.rgba (colour map),
 * OUT_PREFIX.coord (x, y, map u32 triples) and OUT_PREFIX.ids (patch, map
 * u32 pairs) of an OUT_SIDE^2 synthesis with default parameters and the
 * given seed, and OUT_PREFIX.batch.rgba of the same job run through a
 * tsg_batch. Prints one line per check; exits non-zero on any failure.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tsgpu.h"

static int fails = 0;

#define CHECK(cond, what)                                                              \
    do {                                                                               \
        if (cond) {                                                                    \
            printf("ok: %s\n", what);                                                  \
        } else {                                                                       \
            printf("FAIL: %s (%s)\n", what, tsg_last_error());                         \
            ++fails;                                                                   \
        }                                                                              \
    } while (0)

struct progress_seen {
    uint64_t calls, last_total, total;
    int monotone;
};

static void on_progress(void* user, const uint8_t* rgba, uint32_t w, uint32_t h, uint64_t total_current,
                        uint64_t total, uint64_t stage_current, uint64_t stage_total) {
    struct progress_seen* p = (struct progress_seen*)user;
    (void)rgba;
    (void)w;
    (void)h;
    (void)stage_current;
    (void)stage_total;
    if (total_current < p->last_total) p->monotone = 0;
    p->last_total = total_current;
    p->total = total;
    p->calls += 1;
}

static uint8_t* read_file(const char* path, size_t n) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    uint8_t* buf = (uint8_t*)malloc(n);
    size_t got = buf ? fread(buf, 1, n, f) : 0;
    fclose(f);
    if (got != n) {
        free(buf);
        return NULL;
    }
    return buf;
}

static int write_file(const char* prefix, const char* ext, const void* data, size_t n) {
    char path[4096];
    snprintf(path, sizeof path, "%s.%s", prefix, ext);
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    size_t put = fwrite(data, 1, n, f);
    fclose(f);
    return put == n ? 0 : -1;
}

int main(int argc, char** argv) {
    if (argc != 7) {
        fprintf(stderr, "usage: %s EXAMPLE.rgba W H OUT_SIDE SEED OUT_PREFIX\n", argv[0]);
        return 2;
    }
    const uint32_t w = (uint32_t)atoi(argv[2]), h = (uint32_t)atoi(argv[3]);
    const uint32_t side = (uint32_t)atoi(argv[4]);
    const uint64_t seed = (uint64_t)strtoull(argv[5], NULL, 10);
    uint8_t* ex = read_file(argv[1], (size_t)w * h * 4);
    if (!ex) {
        fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    printf("%s\n", tsg_version());
    tsg_ctx* ctx = NULL;
    CHECK(tsg_ctx_create(0, &ctx) == TSG_OK, "tsg_ctx_create");
    if (!ctx) return 1;

    tsg_params p;
    tsg_params_default(&p);
    p.seed = seed;
    tsg_image img = {w, h, ex};

    /* Error::InvalidRange from the builder's validation (session.rs:450-499) */
    tsg_params bad = p;
    bad.cauchy_dispersion = 2.0f;
    tsg_job* job = NULL;
    CHECK(tsg_job_create(ctx, &bad, 1, &img, NULL, NULL, NULL, NULL, NULL, 0, side, side, &job) ==
              TSG_ERR_INVALID_RANGE && job == NULL,
          "cauchy_dispersion 2.0 is TSG_ERR_INVALID_RANGE");

    /* Session::run with a progress callback */
    CHECK(tsg_job_create(ctx, &p, 1, &img, NULL, NULL, NULL, NULL, NULL, 0, side, side, &job) == TSG_OK,
          "tsg_job_create");
    struct progress_seen seen = {0, 0, 0, 1};
    CHECK(tsg_job_run(job, on_progress, &seen) == TSG_OK, "tsg_job_run");
    CHECK(seen.calls > 0 && seen.monotone && seen.last_total == seen.total, "progress callback");
    const size_t n = (size_t)side * side;
    uint8_t* rgba = (uint8_t*)malloc(n * 4);
    uint32_t* coord = (uint32_t*)malloc(n * 3 * 4);
    uint32_t* ids = (uint32_t*)malloc(n * 2 * 4);
    CHECK(tsg_job_download(job, rgba, coord, ids, NULL, NULL) == TSG_OK, "tsg_job_download");
    CHECK(tsg_job_run(job, NULL, NULL) != TSG_OK, "a job runs once");
    tsg_job_destroy(job);
    CHECK(write_file(argv[6], "rgba", rgba, n * 4) == 0 && write_file(argv[6], "coord", coord, n * 12) == 0 &&
              write_file(argv[6], "ids", ids, n * 8) == 0,
          "outputs written");

    /* the same job through the batch API */
    tsg_batch* b = NULL;
    CHECK(tsg_batch_create(ctx, 4, &b) == TSG_OK, "tsg_batch_create");
    job = NULL;
    uint64_t ticket = 0;
    CHECK(tsg_job_create(ctx, &p, 1, &img, NULL, NULL, NULL, NULL, NULL, 0, side, side, &job) == TSG_OK &&
              tsg_batch_submit(b, job, &ticket) == TSG_OK,
          "tsg_batch_submit");
    CHECK(tsg_batch_run_all(b) == TSG_OK, "tsg_batch_run_all");
    uint64_t t_done = 0;
    int status = -1;
    uint32_t n_done = 0;
    CHECK(tsg_batch_poll(b, &t_done, &status, 1, &n_done) == TSG_OK && n_done == 1 && t_done == ticket &&
              status == TSG_OK,
          "tsg_batch_poll");
    tsg_job* back = NULL;
    CHECK(tsg_batch_take(b, ticket, &back) == TSG_OK && back == job, "tsg_batch_take");
    uint8_t* rgba2 = (uint8_t*)malloc(n * 4);
    CHECK(back && tsg_job_download(back, rgba2, NULL, NULL, NULL, NULL) == TSG_OK, "download of the batched job");
    CHECK(memcmp(rgba, rgba2, n * 4) == 0, "batched result equals the run alone");
    CHECK(write_file(argv[6], "batch.rgba", rgba2, n * 4) == 0, "batched output written");
    if (back) tsg_job_destroy(back);
    tsg_batch_destroy(b);
    tsg_ctx_destroy(ctx);
    free(ex);
    free(rgba);
    free(rgba2);
    free(coord);
    free(ids);
    printf("%s\n", fails ? "FAILED" : "ALL OK");
    return fails ? 1 : 0;
}
