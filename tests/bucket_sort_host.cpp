// Host emulation of the bucket sort (gcs_sort.hpp) around its own helper
// functions: split_candidate, bucket_window, bucket_of, sort4, merge_round
// and block_merge run here exactly as the kernels call them, with each
// kernel's thread loop written out, and the result is compared with
// std::stable_sort per instance.  Inputs: nearly sorted keys (the case the
// sort is built for), keys displaced widely and random keys (buckets above
// kBucketCap, merged in chunks), equal keys, and a forest with instances at
// the bucket edges.  Prints one line per case and "ok" at the end; exits
// non-zero on any mismatch or index out of range.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "gcs_sort.hpp"

using namespace gcs;

static long g_fail = 0;

static void check(bool ok, const char* what)
{
    if (!ok && g_fail++ < 10) std::printf("FAIL: %s\n", what);
}

// One bucket of m pairs sorted as k_bucket_sort does: m <= kBucketCap in one
// LDS sort (sort4 per thread, then merge rounds), else in LDS-sized chunks
// merged pairwise by block_merge, the buffers swapping each round.
static void sort_bucket(std::vector<uint64_t>& kin, std::vector<uint32_t>& vin,
                        std::vector<uint64_t>& kout, std::vector<uint32_t>& vout, size_t off, int m)
{
    auto lds_sort = [&](size_t base, int mc) {
        std::vector<uint64_t> sk(kBucketCap + 1, 0);
        std::vector<uint32_t> sv(kBucketCap + 1, 0);
        for (int i = 0; i < kBucketCap; ++i) {
            sk[i] = i < mc ? kin[base + i] : ~0ull;
            sv[i] = i < mc ? vin[base + i] : 0xFFFFFFFFu;
        }
        for (int t = 0; t < kBucketThreads; ++t) sort4(&sk[kBucketPer * t], &sv[kBucketPer * t]);
        for (int w = kBucketPer; w < kBucketCap; w <<= 1) {
            std::vector<uint64_t> nk(kBucketCap + 1, 0);
            std::vector<uint32_t> nv(kBucketCap + 1, 0);
            for (int t = 0; t < kBucketThreads; ++t)
                merge_round(sk.data(), sv.data(), w, kBucketPer * t, &nk[kBucketPer * t],
                            &nv[kBucketPer * t]);
            sk.swap(nk);
            sv.swap(nv);
        }
        for (int i = 0; i < mc; ++i) {
            kout[base + i] = sk[i];
            vout[base + i] = sv[i];
        }
    };
    if (m <= kBucketCap) {
        lds_sort(off, m);
        return;
    }
    for (int c = 0; c < m; c += kBucketCap) lds_sort(off + c, std::min(kBucketCap, m - c));
    uint64_t* ks = kout.data() + off;
    uint32_t* vs = vout.data() + off;
    uint64_t* kd = kin.data() + off;
    uint32_t* vd = vin.data() + off;
    for (int w = kBucketCap; w < m; w <<= 1) {
        for (int s0 = 0; s0 < m; s0 += 2 * w) {
            const int la = std::min(w, m - s0), lb = std::min(w, m - s0 - la);
            for (int t = 0; t < kBucketThreads; ++t)
                block_merge(t, ks + s0, vs + s0, la, ks + s0 + la, vs + s0 + la, lb, kd + s0,
                            vd + s0);
        }
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != kout.data() + off)
        for (int i = 0; i < m; ++i) {
            kout[off + i] = ks[i];
            vout[off + i] = vs[i];
        }
}

static void run(const std::vector<uint64_t>& keys, const std::vector<int>& sizes,
                const char* name)
{
    std::vector<int> start;
    int n_all = 0;
    for (int n : sizes) {
        start.push_back(n_all);
        n_all += n;
    }
    const int k = (int)sizes.size();
    std::vector<int> bkt0(k + 1, 0);
    for (int s = 0; s < k; ++s)
        bkt0[s + 1] = bkt0[s] + std::max(1, (sizes[s] + kBucketKeys - 1) / kBucketKeys);
    const int nbkt = bkt0[k];
    std::vector<uint64_t> split(nbkt);
    std::vector<uint32_t> bcnt(nbkt, 0), boff(nbkt), bfill(nbkt, 0), bid(n_all);

    // k_bucket_cand + k_bucket_split.
    for (int s = 0; s < k; ++s) {
        const int b0 = bkt0[s], nb = bkt0[s + 1] - b0;
        const uint64_t* kk = keys.data() + start[s];
        uint64_t run_max = 0;
        for (int b = 0; b < nb; ++b) {
            uint64_t v = 0;
            if (b > 0) {
                uint64_t x = split_candidate(kk, sizes[s], b - 1, nb);
                uint64_t y = split_candidate(kk, sizes[s], b, nb);
                const uint64_t z = split_candidate(kk, sizes[s], b + 1, nb);
                if (x > y) swap_u64(x, y);
                v = max_u64(x, min_u64(y, z));
            }
            run_max = std::max(run_max, v);
            split[b0 + b] = run_max;
        }
    }
    // k_bucket_count over tiles of kTile keys.
    long outside = 0;
    for (int s = 0; s < k; ++s) {
        const int b0 = bkt0[s], nb = bkt0[s + 1] - b0;
        for (int j0 = 0; j0 < sizes[s]; j0 += kTile) {
            int wbase, wn;
            bucket_window(j0, nb, wbase, wn);
            check(wbase >= 0 && wn >= 1 && wbase + wn <= nb, "window inside the instance");
            uint64_t ls[kBucketWin + 1];
            for (int w = 0; w <= wn; ++w) ls[w] = wbase + w < nb ? split[b0 + wbase + w] : ~0ull;
            for (int i = j0; i < std::min(sizes[s], j0 + kTile); ++i) {
                const uint64_t key = keys[start[s] + i];
                bool inwin;
                const int b = bucket_of(key, ls, wbase, wn, split.data() + b0, nb, inwin);
                int brute = 0;
                for (int q = 1; q < nb; ++q) brute += split[b0 + q] <= key;
                check(b == brute, "bucket equals the count of splitters at most the key");
                check(!inwin || (b >= wbase && b < wbase + wn), "window bucket inside the window");
                outside += !inwin;
                bcnt[b0 + b]++;
                bid[start[s] + i] = (uint32_t)(b0 + b);
            }
        }
    }
    // k_bucket_scan.
    for (int s = 0; s < k; ++s) {
        uint32_t c = (uint32_t)start[s];
        for (int b = bkt0[s]; b < bkt0[s + 1]; ++b) {
            boff[b] = c;
            c += bcnt[b];
        }
    }
    // k_bucket_scatter, in an arbitrary order within each bucket.
    std::vector<uint64_t> k1(n_all), k0(n_all);
    std::vector<uint32_t> v1(n_all), v0(n_all);
    std::vector<int> order(n_all);
    for (int i = 0; i < n_all; ++i) order[i] = i;
    std::shuffle(order.begin(), order.end(), std::mt19937(5));
    for (int i : order) {
        const uint32_t b = bid[i];
        const uint32_t pos = boff[b] + bfill[b]++;
        check(pos < (uint32_t)n_all, "scatter inside the arrays");
        if (pos >= (uint32_t)n_all) return;
        k1[pos] = keys[i];
        v1[pos] = (uint32_t)i;
    }
    // k_bucket_sort.
    int largest = 0, chunked = 0;
    for (int b = 0; b < nbkt; ++b) {
        largest = std::max(largest, (int)bcnt[b]);
        chunked += bcnt[b] > (uint32_t)kBucketCap;
        if (bcnt[b]) sort_bucket(k1, v1, k0, v0, boff[b], (int)bcnt[b]);
    }
    long bad = 0;
    for (int s = 0; s < k; ++s) {
        std::vector<uint32_t> idx(sizes[s]);
        for (int i = 0; i < sizes[s]; ++i) idx[i] = (uint32_t)(start[s] + i);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t c) { return keys[a] < keys[c]; });
        for (int i = 0; i < sizes[s]; ++i) {
            const size_t g = (size_t)start[s] + i;
            bad += v0[g] != idx[i] || k0[g] != keys[idx[i]];
        }
    }
    check(bad == 0, name);
    std::printf("%-26s n=%d buckets=%d largest=%d chunked=%d outside-window=%ld mismatches=%ld\n",
                name, n_all, nbkt, largest, chunked, outside, bad);
}

int main()
{
    std::mt19937_64 r(1);
    auto random_keys = [&](int n) {
        std::vector<uint64_t> k(n);
        for (auto& x : k) x = r() >> 1;  // Morton keys have 63 bits
        return k;
    };
    // Sorted keys, each swapped with one up to `disp` places away, and a
    // fraction `far` replaced by random keys (the jumps across the key space).
    auto nearly_sorted = [&](int n, int disp, double far) {
        std::vector<uint64_t> k = random_keys(n);
        std::sort(k.begin(), k.end());
        for (int i = 0; i < n; ++i) {
            const int j = std::min(n - 1, std::max(0, i + (int)(r() % (2 * disp + 1)) - disp));
            std::swap(k[i], k[j]);
        }
        for (int i = 0; i < n; ++i)
            if ((double)(r() % 1000000) < far * 1e6) k[i] = r() >> 1;
        return k;
    };
    run(nearly_sorted(300000, 200, 0.001), {300000}, "nearly sorted");
    run(nearly_sorted(120000, 5000, 0.01), {120000}, "displaced widely");
    run(random_keys(50000), {50000}, "random order");
    run(std::vector<uint64_t>(30000, 12345), {30000}, "all keys equal");
    {
        std::vector<uint64_t> k = nearly_sorted(40000, 50, 0.0);
        for (int i = 10000; i < 30000; ++i) k[i] = k[10000];
        run(k, {40000}, "a run of equal keys");
    }
    {
        std::vector<uint64_t> k(40000);
        for (auto& x : k) x = r() % 3;
        run(k, {40000}, "three distinct keys");
    }
    {
        const std::vector<int> sizes = {1, 2, 3, 1535, 1536, 1537, 3072, 3073, 1023, 1024, 1025, 20000};
        std::vector<uint64_t> k;
        for (int n : sizes) {
            const std::vector<uint64_t> part = nearly_sorted(n, 100, 0.002);
            k.insert(k.end(), part.begin(), part.end());
        }
        run(k, sizes, "ragged forest");
    }
    std::printf("%s\n", g_fail ? "FAIL" : "ok");
    return g_fail ? 1 : 0;
}
