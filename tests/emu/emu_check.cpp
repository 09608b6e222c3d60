// Runs the library's C ABI, built from the unmodified HIP sources on the CPU
// emulation (hip_emu.hpp), under AddressSanitizer + UBSan, over the shapes the
// GPU tests use, and compares every result with the C oracle.  Exit status 0
// when every case matches; ASan aborts on any out-of-bounds access.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "gpumultisplit.h"

extern "C" {
int ms_oracle_multisplit(uint32_t kind, uint32_t B, uint32_t param, const uint32_t* kin,
                         const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n,
                         uint32_t* offsets);
int ms_oracle_sort(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                   uint64_t n, uint32_t digit_bits);
void ms_oracle_histogram(uint32_t kind, uint32_t B, uint32_t param, const uint32_t* keys,
                         uint64_t n, uint32_t* hist);
uint64_t ms_oracle_compaction(uint32_t kind, uint32_t param, const uint32_t* kin,
                              const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n);
void ms_oracle_histogram_range_f32(const float* x, uint64_t n, int32_t* hist, const float* levels,
                                   uint32_t nl);
void ms_oracle_histogram_range_u32(const uint32_t* x, uint64_t n, int32_t* hist,
                                   const uint32_t* levels, uint32_t nl);
int ms_oracle_multisplit_wide(uint32_t kind, uint32_t B, uint64_t param, uint32_t width,
                              uint32_t key_words, uint32_t value_words, const uint32_t* kin,
                              const uint32_t* vin, uint32_t* kout, uint32_t* vout, uint64_t n,
                              uint32_t* offsets);
int ms_oracle_sort_wide(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                        uint64_t n, uint32_t value_words, uint32_t digit_bits);
void ms_oracle_tile_histograms(uint32_t kind, uint32_t B, uint32_t param, const uint32_t* keys,
                               uint64_t n, uint32_t tile, uint32_t* H);
}

static uint64_t rng_state = 0x1234567;
static uint32_t rnd() {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

static int failures = 0;
#define EXPECT(cond, ...)                   \
    do {                                    \
        if (!(cond)) {                      \
            printf("FAIL: " __VA_ARGS__);   \
            printf("\n");                   \
            ++failures;                     \
        }                                   \
    } while (0)

// Exact-size heap buffers so that ASan sees any access past n.
struct Buf {
    uint32_t* p;
    size_t n;
    explicit Buf(size_t n_, uint32_t fill = 0x5A5A5A5Au) : n(n_) {
        p = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
        for (size_t i = 0; i < n; ++i) p[i] = fill;
    }
    ~Buf() { free(p); }
    Buf(const Buf&) = delete;
};

static void multisplit_case(uint32_t B, uint32_t n, gms_bucket_desc d, uint32_t oracle_param,
                            bool kv, uint32_t key_mod, int misalign) {
    Buf kin(n + 4), vin(n + 4), kout(n + 4), vout(n + 4);
    for (uint32_t i = 0; i < n + 4; ++i) {
        kin.p[i] = key_mod ? rnd() % key_mod : rnd();
        vin.p[i] = i;
    }
    const uint32_t* ki = kin.p + misalign;
    const uint32_t* vi = vin.p + (misalign ? 3 : 0);
    uint32_t* ko = kout.p + (misalign ? 2 : 0);
    uint32_t* vo = vout.p + (misalign ? 1 : 0);
    gms_multisplit_ctx* ctx = nullptr;
    EXPECT(gms_multisplit_create(B, &ctx) == GMS_SUCCESS, "create B=%u", B);
    EXPECT(gms_multisplit_allocate(ctx, n, kv) == GMS_SUCCESS, "allocate");
    std::vector<uint32_t> offs(B, 77);
    gms_status st = kv ? gms_multisplit_key_value(ctx, ki, vi, ko, vo, n, d, offs.data(), nullptr)
                       : gms_multisplit_key_only(ctx, ki, ko, n, d, offs.data(), nullptr);
    EXPECT(st == GMS_SUCCESS, "multisplit status %d", (int)st);
    std::vector<uint32_t> wk(n + 1), wv(n + 1), wo(B);
    ms_oracle_multisplit(d.kind, B, oracle_param, ki, kv ? vi : nullptr, wk.data(),
                         kv ? wv.data() : nullptr, n, wo.data());
    bool ok = memcmp(ko, wk.data(), n * 4) == 0 && memcmp(offs.data(), wo.data(), B * 4) == 0 &&
              (!kv || memcmp(vo, wv.data(), n * 4) == 0);
    EXPECT(ok, "multisplit kind=%u B=%u n=%u kv=%d misalign=%d", d.kind, B, n, (int)kv, misalign);
    gms_multisplit_destroy(ctx);
}

int main(int argc, char** argv) {
    const bool quick = argc > 1 && strcmp(argv[1], "quick") == 0;
    const uint32_t tile_kv = gms_subproblem_size(256, 1);
    const uint32_t sizes[] = {0, 1, 2, 63, 64, 65, 1000, tile_kv - 1, tile_kv, tile_kv + 1,
                              3 * tile_kv + 17};
    const uint32_t bs[] = {1, 2, 3, 4, 5, 8, 16, 17, 32, 33, 64, 100, 128, 129, 255, 256,
                           300, 512, 1000, 1024};
    for (uint32_t B : bs) {
        for (uint32_t n : sizes) {
            if (quick && n > tile_kv + 1) continue;
            for (int kv = 0; kv < 2; ++kv)
                multisplit_case(B, n, gms_bucket_desc{GMS_BUCKET_RANGE, 0, 0}, 0, kv, 0, 0);
        }
        printf("range B=%u done\n", B);
        fflush(stdout);
    }
    // other identifiers
    multisplit_case(32, 20000, gms_bucket_desc{GMS_BUCKET_DELTA, 625, 0}, 625, true, 20000, 0);
    for (uint32_t delta : {1u, 3u, 1000u, 0x01000001u, 0xFFFFFFFFu})
        multisplit_case(7, 9000, gms_bucket_desc{GMS_BUCKET_DELTA, delta, 0}, delta, true, 0, 0);
    multisplit_case(16, 9000, gms_bucket_desc{GMS_BUCKET_IDENTITY, 0, 0}, 0, true, 16, 0);
    multisplit_case(16, 9000, gms_bucket_desc{GMS_BUCKET_IDENTITY, 0, 0}, 0, true, 32, 0);
    multisplit_case(256, 9000, gms_bucket_desc{GMS_BUCKET_DIGIT, 24, 8}, 24, true, 0, 0);
    multisplit_case(128, 9000, gms_bucket_desc{GMS_BUCKET_DIGIT, 25, 7}, 25, true, 0, 0);
    multisplit_case(8, 9000, gms_bucket_desc{GMS_BUCKET_LOWBITS, 3, 0}, 3, false, 0, 0);
    multisplit_case(128, 2 * tile_kv + 5, gms_bucket_desc{GMS_BUCKET_RANGE, 0, 0}, 0, true, 0, 1);
    multisplit_case(256, 2 * tile_kv, gms_bucket_desc{GMS_BUCKET_RANGE, 0, 0}, 0, false, 0, 1);
    printf("identifiers done\n");
    fflush(stdout);

    // prescan matrix and scan stage
    {
        const uint32_t n = 3 * tile_kv + 5, B = 256;
        Buf keys(n);
        for (uint32_t i = 0; i < n; ++i) keys.p[i] = rnd();
        gms_multisplit_ctx* ctx;
        gms_multisplit_create(B, &ctx);
        gms_multisplit_allocate(ctx, n, 1);
        const uint32_t nt = (n + tile_kv - 1) / tile_kv;
        Buf mat(B * nt);
        std::vector<uint32_t> want(B * nt);
        EXPECT(gms_multisplit_prescan_matrix(ctx, keys.p, n, gms_bucket_desc{0, 0, 0}, 1, mat.p,
                                             nullptr) == GMS_SUCCESS,
               "prescan matrix status");
        ms_oracle_tile_histograms(0, B, 0, keys.p, n, tile_kv, want.data());
        EXPECT(memcmp(mat.p, want.data(), B * nt * 4) == 0, "prescan matrix");
        for (uint32_t m : {1u, 5u, 4095u, 4096u, 4097u, 20000u}) {
            if (m > B * nt) continue;
            Buf in(m), out(m);
            uint64_t run = 0;
            std::vector<uint32_t> ex(m);
            for (uint32_t i = 0; i < m; ++i) {
                in.p[i] = rnd() % 1000;
                ex[i] = (uint32_t)run;
                run += in.p[i];
            }
            EXPECT(gms_exclusive_scan_u32(ctx, in.p, out.p, m, nullptr) == GMS_SUCCESS, "scan st");
            EXPECT(memcmp(out.p, ex.data(), m * 4) == 0, "scan m=%u", m);
            EXPECT(gms_exclusive_scan_u32(ctx, in.p, in.p, m, nullptr) == GMS_SUCCESS, "scan st");
            EXPECT(memcmp(in.p, ex.data(), m * 4) == 0, "scan in place m=%u", m);
        }
        gms_multisplit_destroy(ctx);
        printf("stages done\n");
        fflush(stdout);
    }

    // sort
    for (uint32_t bits : {4u, 5u, 6u, 7u, 8u}) {
        for (uint32_t n : {0u, 1u, 1000u, tile_kv + 3}) {
            for (int kv = 0; kv < 2; ++kv) {
                Buf kin(n), vin(n), kout(n), vout(n);
                for (uint32_t i = 0; i < n; ++i) {
                    kin.p[i] = (i % 7 == 0) ? 12345u : rnd();
                    vin.p[i] = i;
                }
                gms_sort_ctx* ctx;
                gms_sort_create(&ctx);
                EXPECT(gms_sort_allocate(ctx, n, bits, kv) == GMS_SUCCESS, "sort alloc");
                EXPECT(gms_sort(ctx, kin.p, kout.p, kv ? vin.p : nullptr, kv ? vout.p : nullptr, n,
                                bits, nullptr) == GMS_SUCCESS,
                       "sort status");
                std::vector<uint32_t> wk(n + 1), wv(n + 1);
                ms_oracle_sort(kin.p, kv ? vin.p : nullptr, wk.data(), kv ? wv.data() : nullptr,
                               n, bits);
                EXPECT(memcmp(kout.p, wk.data(), n * 4) == 0 &&
                           (!kv || memcmp(vout.p, wv.data(), n * 4) == 0),
                       "sort bits=%u n=%u kv=%d", bits, n, kv);
                gms_sort_destroy(ctx);
            }
        }
    }
    printf("sort done\n");
    fflush(stdout);

    // compaction and histogram
    for (uint32_t n : {0u, 1u, 5000u, tile_kv + 9}) {
        Buf kin(n), vin(n), kout(n), vout(n);
        for (uint32_t i = 0; i < n; ++i) {
            kin.p[i] = rnd();
            vin.p[i] = i;
        }
        gms_compaction_ctx* c;
        gms_compaction_create(&c);
        gms_compaction_allocate(c, n, 1);
        uint32_t cnt = 99;
        EXPECT(gms_compaction(c, kin.p, vin.p, kout.p, vout.p, n,
                              gms_bucket_desc{GMS_BUCKET_LOWBITS, 1, 0}, &cnt, nullptr,
                              nullptr) == GMS_SUCCESS,
               "compaction status");
        std::vector<uint32_t> wk(n + 1), wv(n + 1);
        uint64_t wc = ms_oracle_compaction(GMS_BUCKET_LOWBITS, 1, kin.p, vin.p, wk.data(),
                                           wv.data(), n);
        EXPECT(cnt == wc && memcmp(kout.p, wk.data(), n * 4) == 0 &&
                   memcmp(vout.p, wv.data(), n * 4) == 0,
               "compaction n=%u", n);
        gms_compaction_destroy(c);

        Buf hist(256, 0);
        std::vector<uint32_t> wh(256, 0);
        EXPECT(gms_histogram_u32(kin.p, n, hist.p, 256, gms_bucket_desc{GMS_BUCKET_DIGIT, 0, 8},
                                 nullptr) == GMS_SUCCESS,
               "histogram status");
        ms_oracle_histogram(GMS_BUCKET_DIGIT, 256, 0, kin.p, n, wh.data());
        EXPECT(memcmp(hist.p, wh.data(), 256 * 4) == 0, "histogram n=%u", n);

        // custom bin edges: 33 and 257 levels, integer and float samples
        for (uint32_t nl : {2u, 33u, 257u}) {
            Buf lv(nl), hr(nl - 1, 0);
            for (uint32_t i = 0; i < nl; ++i) lv.p[i] = 1000u + i * (0xF0000000u / nl);
            std::vector<int32_t> wr(nl - 1, 0);
            EXPECT(gms_histogram_range_u32(kin.p, n, (int32_t*)hr.p, nl, lv.p, nullptr) ==
                       GMS_SUCCESS,
                   "histogram_range_u32 status");
            ms_oracle_histogram_range_u32(kin.p, n, wr.data(), lv.p, nl);
            EXPECT(memcmp(hr.p, wr.data(), (nl - 1) * 4) == 0, "histogram_range_u32 n=%u nl=%u", n,
                   nl);
            std::vector<float> xs(n ? n : 1), fl(nl);
            for (uint32_t i = 0; i < n; ++i) xs[i] = (float)(kin.p[i] >> 8) * (1.0f / 65536.0f) - 8.0f;
            if (n > 3) xs[3] = __builtin_nanf("");
            for (uint32_t i = 0; i < nl; ++i) fl[i] = -6.0f + 240.0f * (float)i / (float)nl;
            Buf hf(nl - 1, 0);
            std::vector<int32_t> wf(nl - 1, 0);
            EXPECT(gms_histogram_range_f32(xs.data(), n, (int32_t*)hf.p, nl, fl.data(), nullptr) ==
                       GMS_SUCCESS,
                   "histogram_range_f32 status");
            ms_oracle_histogram_range_f32(xs.data(), n, wf.data(), fl.data(), nl);
            EXPECT(memcmp(hf.p, wf.data(), (nl - 1) * 4) == 0, "histogram_range_f32 n=%u nl=%u", n,
                   nl);
        }
    }
    printf("compaction/histogram done\n");

    // batched multisplit: exact-size buffers per instance, every tuning
    for (uint32_t K : {8u, 16u, 24u, 32u}) {
        for (uint32_t pers : {0u, 1u}) {
            const uint32_t ns[] = {0, 1, 777, 2048, 4096 + 3, 6144, 13000};
            const uint32_t nbs[] = {8, 2, 1, 100, 64, 256, 17};
            const int nk = 7;
            std::vector<Buf*> bufs;
            std::vector<gms_batch_item> items;
            for (int i = 0; i < nk; ++i) {
                const uint32_t n = ns[i], B = nbs[i];
                const bool kv = i % 2 == 0;
                Buf *ki = new Buf(n), *vi = new Buf(n), *ko = new Buf(n), *vo = new Buf(n),
                    *of = new Buf(B);
                for (uint32_t j = 0; j < n; ++j) {
                    ki->p[j] = rnd();
                    vi->p[j] = j;
                }
                bufs.insert(bufs.end(), {ki, vi, ko, vo, of});
                items.push_back(gms_batch_item{ki->p, kv ? vi->p : nullptr, ko->p,
                                               kv ? vo->p : nullptr, n, B,
                                               gms_bucket_desc{GMS_BUCKET_RANGE, 0, 0}, of->p});
            }
            gms_batch* b;
            gms_batch_create(&b);
            EXPECT(gms_batch_set_tuning(b, gms_batch_tuning{K, pers}) == GMS_SUCCESS, "tuning");
            EXPECT(gms_batch_plan(b, items.data(), nk, nullptr) == GMS_SUCCESS, "batch plan");
            EXPECT(gms_batch_run(b, nullptr) == GMS_SUCCESS, "batch run");
            gms_batch_destroy(b);
            for (int i = 0; i < nk; ++i) {
                const gms_batch_item& it = items[i];
                std::vector<uint32_t> wk(it.n + 1), wv(it.n + 1), wo(it.num_buckets);
                ms_oracle_multisplit(GMS_BUCKET_RANGE, it.num_buckets, 0, it.d_key_in,
                                     it.d_value_in, wk.data(), it.d_value_in ? wv.data() : nullptr,
                                     it.n, wo.data());
                EXPECT(memcmp(it.d_key_out, wk.data(), it.n * 4) == 0 &&
                           (!it.d_value_in || memcmp(it.d_value_out, wv.data(), it.n * 4) == 0) &&
                           memcmp(it.d_offsets, wo.data(), it.num_buckets * 4) == 0,
                       "batch K=%u persistent=%u instance %d", K, pers, i);
            }
            for (Buf* x : bufs) delete x;
        }
    }
    printf("batch done\n");

    // 64-bit keys / values: exact-size buffers (ASan sees any overrun)
    for (uint32_t kw : {1u, 2u}) {
        for (uint32_t vw : {0u, 1u, 2u}) {
            if (kw == 1 && vw < 2) continue;
            for (uint32_t n : {0u, 1u, 3072u * 2 + 5, 4096u + 1}) {
                const uint32_t B = 256;
                Buf kin(n * kw), vin(n * vw), kout(n * kw), vout(n * vw);
                for (uint32_t i = 0; i < n * kw; ++i) kin.p[i] = rnd();
                for (uint32_t i = 0; i < n * vw; ++i) vin.p[i] = i;
                gms_multisplit_ctx* c;
                gms_multisplit_create(B, &c);
                gms_multisplit_allocate(c, n, 1);
                std::vector<uint32_t> offs(B), wk(n * kw + 1), wv(n * vw + 1), wo(B);
                EXPECT(gms_multisplit_wide(c, kin.p, vw ? vin.p : nullptr, kout.p,
                                           vw ? vout.p : nullptr, n, 4 * kw, 4 * vw,
                                           gms_bucket_desc64{GMS_BUCKET_RANGE, 0, 0}, nullptr,
                                           offs.data(), nullptr) == GMS_SUCCESS,
                       "wide status kw=%u vw=%u n=%u", kw, vw, n);
                ms_oracle_multisplit_wide(0, B, 0, 0, kw, vw, kin.p, vin.p, wk.data(), wv.data(), n,
                                          wo.data());
                EXPECT(memcmp(kout.p, wk.data(), (size_t)n * kw * 4) == 0 &&
                           memcmp(vout.p, wv.data(), (size_t)n * vw * 4) == 0 && offs == wo,
                       "wide multisplit kw=%u vw=%u n=%u", kw, vw, n);
                gms_multisplit_destroy(c);
                if (kw == 2) {
                    gms_sort_ctx* sc;
                    gms_sort_create(&sc);
                    EXPECT(gms_sort_allocate_wide(sc, n, 8, 8, 4 * vw) == GMS_SUCCESS, "sort alloc");
                    EXPECT(gms_sort_wide(sc, kin.p, kout.p, vw ? vin.p : nullptr,
                                         vw ? vout.p : nullptr, n, 8, 4 * vw, nullptr) == GMS_SUCCESS,
                           "sort wide status");
                    ms_oracle_sort_wide(kin.p, vin.p, wk.data(), wv.data(), n, vw, 8);
                    EXPECT(memcmp(kout.p, wk.data(), (size_t)n * 8) == 0 &&
                               memcmp(vout.p, wv.data(), (size_t)n * vw * 4) == 0,
                           "wide sort vw=%u n=%u", vw, n);
                    gms_sort_destroy(sc);
                }
            }
        }
    }
    // 64-bit elements into more than 256 buckets on one context whose
    // capacity grows between calls: the pass temporaries of the first call
    // must not be reused at the larger size (ASan sees an overrun if they are).
    {
        const uint32_t B = 1024, n1 = 1000, n2 = 3072u * 3 + 11;
        gms_multisplit_ctx* c;
        gms_multisplit_create(B, &c);
        for (uint32_t n : {n1, n2}) {
            EXPECT(gms_multisplit_allocate(c, n, 1) == GMS_SUCCESS, "allocate %u", n);
            Buf kin(n * 2), vin(n * 2), kout(n * 2), vout(n * 2);
            for (uint32_t i = 0; i < n * 2; ++i) {
                kin.p[i] = rnd();
                vin.p[i] = i;
            }
            std::vector<uint32_t> offs(B), wk(n * 2), wv(n * 2), wo(B);
            EXPECT(gms_multisplit_wide(c, kin.p, vin.p, kout.p, vout.p, n, 8, 8,
                                       gms_bucket_desc64{GMS_BUCKET_RANGE, 0, 0}, nullptr,
                                       offs.data(), nullptr) == GMS_SUCCESS,
                   "wide regrow status n=%u", n);
            ms_oracle_multisplit_wide(0, B, 0, 0, 2, 2, kin.p, vin.p, wk.data(), wv.data(), n,
                                      wo.data());
            EXPECT(memcmp(kout.p, wk.data(), (size_t)n * 8) == 0 &&
                       memcmp(vout.p, wv.data(), (size_t)n * 8) == 0 && offs == wo,
                   "wide regrow n=%u", n);
        }
        gms_multisplit_destroy(c);
    }
    printf("wide done\n");

    // wave-level variant: exact-size buffers, ragged subproblems
    for (uint32_t B : {2u, 5u, 64u}) {
        for (uint32_t n : {1u, 1023u, 4096u * 2 + 200}) {
            for (int kv = 0; kv < 2; ++kv) {
                Buf kin(n), vin(n), kout(n), vout(n);
                for (uint32_t i = 0; i < n; ++i) {
                    kin.p[i] = rnd();
                    vin.p[i] = i;
                }
                gms_multisplit_ctx* c;
                gms_multisplit_create(B, &c);
                gms_multisplit_allocate(c, n, kv);
                EXPECT(gms_multisplit_set_variant(c, GMS_VARIANT_WMS) == GMS_SUCCESS, "wms variant");
                std::vector<uint32_t> offs(B), wk(n), wv(n), wo(B);
                const gms_bucket_desc d{GMS_BUCKET_RANGE, 0, 0};
                gms_status st = kv ? gms_multisplit_key_value(c, kin.p, vin.p, kout.p, vout.p, n, d,
                                                              offs.data(), nullptr)
                                   : gms_multisplit_key_only(c, kin.p, kout.p, n, d, offs.data(),
                                                             nullptr);
                EXPECT(st == GMS_SUCCESS, "wms status B=%u n=%u", B, n);
                ms_oracle_multisplit(0, B, 0, kin.p, kv ? vin.p : nullptr, wk.data(),
                                     kv ? wv.data() : nullptr, n, wo.data());
                EXPECT(memcmp(kout.p, wk.data(), (size_t)n * 4) == 0 && offs == wo &&
                           (!kv || memcmp(vout.p, wv.data(), (size_t)n * 4) == 0),
                       "wms B=%u n=%u kv=%d", B, n, kv);
                gms_multisplit_destroy(c);
            }
        }
    }
    printf("wms done\n");
    printf(failures ? "EMU CHECK FAILED (%d)\n" : "EMU CHECK OK\n", failures);
    return failures ? 1 : 0;
}
