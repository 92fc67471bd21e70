// Known-answer harness over the reference's own SYCL-free device primitives.
// TEST INFRASTRUCTURE: built by oracle/ref.mk from the headers where they lie
// under the reference tree; prints JSON vectors that tests/test_oracle.py
// feeds through the oracle (tests/golden/ref_vectors.json keeps a copy so the
// test also runs where the reference tree is absent).
#include <cstdint>
#include <cstdio>
#include <vector>
#include "crystal/load.h"
#include "crystal/pred.h"
#include "crystal/utils.h"

static constexpr int BT = 128, IPT = 4, TILE = BT * IPT;

static uint64_t rng_state = 0x1234567ULL;
static uint32_t rnd() {
    rng_state = rng_state * 6364136223846793005ULL + 1442695040888963407ULL;
    return (uint32_t)(rng_state >> 33);
}

// Run a block predicate chain over one tile: returns flags for tile items.
template <typename F>
static std::vector<int> over_tile(const std::vector<int>& col, int num_items, F body) {
    std::vector<int> out(TILE, -1);
    for (int tid = 0; tid < BT; ++tid) {
        int items[IPT];
        int flags[IPT];
        InitFlags<BT, IPT>(flags);
        BlockLoad<int, BT, IPT>(const_cast<int*>(col.data()), tid, 0, items, num_items);
        body(tid, items, flags, num_items);
        for (int i = 0; i < IPT; ++i) out[tid + i * BT] = flags[i];
    }
    return out;
}

static void print_vec(const char* name, const std::vector<int>& v, int n, bool last = false) {
    std::printf("  \"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%d", i ? "," : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

int main() {
    std::vector<int> col(TILE), col2(TILE);
    for (int i = 0; i < TILE; ++i) { col[i] = (int)(rnd() % 11); col2[i] = 1 + (int)(rnd() % 50); }
    const int n_full = TILE, n_tail = 300;
    std::printf("{\n");
    print_vec("discount", col, TILE);
    print_vec("quantity", col2, TILE);
    // lo_discount >= 1 AND lo_discount <= 3 (BlockPredGTE then BlockPredALTE)
    for (int n : {n_full, n_tail}) {
        auto f = over_tile(col, n, [](int tid, int (&items)[IPT], int (&flags)[IPT], int ni) {
            BlockPredGTE<int, BT, IPT>(tid, items, flags, 1, ni);
            BlockPredALTE<int, BT, IPT>(tid, items, flags, 3, ni);
        });
        char name[64];
        std::snprintf(name, sizeof name, "disc_ge1_le3_n%d", n);
        print_vec(name, f, n);
    }
    // d = 2 OR d = 7 OR d > 9 with a zero-initialised OR accumulator (build-filter form)
    {
        auto f = over_tile(col, n_tail, [](int tid, int (&items)[IPT], int (&flags)[IPT], int ni) {
            int m[IPT];
            InitFlagsZero<BT, IPT>(m);
            BlockPredOrEq<int, BT, IPT>(tid, items, m, 2, ni);
            BlockPredOrEq<int, BT, IPT>(tid, items, m, 7, ni);
            BlockPredOGT<int, BT, IPT>(tid, items, m, 9, ni);
            BlockApplyMaskAnd<BT, IPT>(tid, flags, m);
        });
        print_vec("disc_eq2_or_eq7_or_gt9_n300", f, n_tail);
    }
    // NOT-equal and strict comparisons on quantity
    {
        auto f = over_tile(col2, n_full, [](int tid, int (&items)[IPT], int (&flags)[IPT], int ni) {
            BlockPredNEq<int, BT, IPT>(tid, items, flags, 25, ni);
            BlockPredALT<int, BT, IPT>(tid, items, flags, 40, ni);
            BlockPredAGT<int, BT, IPT>(tid, items, flags, 10, ni);
        });
        print_vec("qty_ne25_lt40_gt10", f, n_full);
    }
    // validity: LSB-first u64 words, BlockLoadValidity + BlockApplyValidityAnd
    {
        std::vector<uint64_t> bitmap((TILE + 63) / 64);
        for (auto& w : bitmap) w = ((uint64_t)rnd() << 32) | rnd();
        std::printf("  \"bitmap\": [");
        for (size_t i = 0; i < bitmap.size(); ++i) std::printf("%s\"%llu\"", i ? "," : "", (unsigned long long)bitmap[i]);
        std::printf("],\n");
        std::vector<int> out(TILE, -1);
        for (int tid = 0; tid < BT; ++tid) {
            int valid[IPT];
            int flags[IPT];
            InitFlags<BT, IPT>(flags);
            BlockLoadValidity<BT, IPT>(bitmap.data(), tid, 0, valid, n_tail);
            BlockApplyValidityAnd<BT, IPT>(tid, flags, valid, n_tail);
            for (int i = 0; i < IPT; ++i) out[tid + i * BT] = flags[i];
        }
        print_vec("validity_and_n300", out, n_tail);
    }
    // utils.h validity decoders
    {
        std::vector<unsigned char> bytes = {1, 0, 0, 1, 1, 0, 1};
        auto v8 = decodeValidityByteBitmap(bytes, bytes.size(), true);
        auto n8 = decodeValidityByteBitmap(bytes, bytes.size(), false);
        auto n64 = decodeValidityWordBitmap(std::vector<uint64_t>{0x5ULL}, 5, false);
        std::printf("  \"valid8_1001101\": \"%llu\",\n", (unsigned long long)v8[0]);
        std::printf("  \"null8_1001101\": \"%llu\",\n", (unsigned long long)n8[0]);
        std::printf("  \"null64_0x5_n5\": \"%llu\"\n", (unsigned long long)n64[0]);
    }
    std::printf("}\n");
    return 0;
}
