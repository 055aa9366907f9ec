// Host-only check of the MFMA assign's work partition (kmpartition.hpp), built
// with g++ by test_partition.py.  Random levels with tasks of every size,
// including tasks whose tiles are not streamed (0 streamed tiles between
// streamed ones), against the invariants the assign and update kernels rely on:
// every streamed tile of every task is covered exactly once, chunk tile indices
// are global and stay inside their task's tiles, a task's chunks are
// contiguous and listed by taskChunkBeg/taskChunkCnt, and each wave gets an
// equal share of the streamed tiles (to within one).  Partials: chunks share
// one exactly when they are consecutive chunks of one task in one workgroup
// (then through an LDS slot of that workgroup, at most wavesPerBlock - 1 per
// workgroup), partials are numbered in chunk order, and taskPartBeg/Cnt list
// each task's partials.  Then a simulation of one assign pass's partial-sum
// bookkeeping (below): the current layout, one tile range per wave, must never
// lose a tile's contribution; the layout of a reverted experiment, one range
// per workgroup whose waves take the tiles in turn, must be caught losing them
// (chunks without a merge slot are written with plain stores, which is only
// sound when one wave processes all of the chunk's tiles).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "kmpartition.hpp"

using namespace rbow_amd;

static int fails = 0;
#define CHECK(c, ...)                                      \
    do {                                                   \
        if (!(c)) {                                        \
            if (fails++ < 20) {                            \
                std::fprintf(stderr, "FAIL %s: ", #c);     \
                std::fprintf(stderr, __VA_ARGS__);         \
                std::fprintf(stderr, "\n");                \
            }                                              \
        }                                                  \
    } while (0)

static void checkParts(const KmPartition& p, uint32_t nbTasks, uint32_t wpb, uint32_t spb, int id) {
    const uint32_t W = std::max(wpb, 1u), S = std::min(spb, W - 1), nbBlocks = (p.nbWaves + W - 1) / W;
    CHECK(p.wavesPerBlock == W && p.slotsPerBlock == S, "case %d: wavesPerBlock / slotsPerBlock", id);
    CHECK(p.blockSlots.size() == size_t(nbBlocks) * S, "case %d: blockSlots size", id);
    std::vector<uint32_t> blockOf(p.chunks.size(), 0u);
    for (uint32_t w = 0; w < p.nbWaves; w++)
        for (uint32_t c = p.waveChunkBeg[w]; c < p.waveChunkBeg[w + 1]; c++) blockOf[c] = w / W;
    // runs of same-task chunks within a workgroup, numbered per workgroup in order:
    // the first S runs share a part through their slot, later ones keep a part per chunk
    std::vector<uint32_t> slotUses(p.blockSlots.size(), 0u);
    uint32_t runsInBlock = 0;
    for (size_t c = 0; c < p.chunks.size(); c++) {
        const KmChunk& k = p.chunks[c];
        CHECK(k.part < p.nbParts, "case %d: chunk %zu part %u of %u", id, c, k.part, p.nbParts);
        if (c == 0) CHECK(k.part == 0, "case %d: first part %u", id, k.part);
        if (c == 0 || blockOf[c] != blockOf[c - 1]) runsInBlock = 0;
        const bool sameRun = c > 0 && blockOf[c] == blockOf[c - 1] && p.chunks[c - 1].task == k.task;
        const bool runNext = c + 1 < p.chunks.size() && blockOf[c + 1] == blockOf[c] && p.chunks[c + 1].task == k.task;
        if (!sameRun && runNext) runsInBlock++;  // chunk c starts a run of two or more
        const bool slotted = (sameRun || runNext) && runsInBlock <= S;
        if (c > 0) {
            if (sameRun && slotted) CHECK(k.part == p.chunks[c - 1].part, "case %d: chunk %zu not merged with its run", id, c);
            else CHECK(k.part == p.chunks[c - 1].part + 1, "case %d: chunk %zu part %u after %u", id, c, k.part,
                       p.chunks[c - 1].part);
        }
        if (slotted) {
            CHECK(k.slot < S, "case %d: merged chunk %zu slot %u", id, c, k.slot);
            if (k.slot < S) {
                const KmSlot& sl = p.blockSlots[size_t(blockOf[c]) * S + k.slot];
                CHECK(sl.part == k.part && sl.task == k.task, "case %d: chunk %zu slot record", id, c);
                slotUses[size_t(blockOf[c]) * S + k.slot]++;
            }
        } else {
            CHECK(k.slot == kNoSlot, "case %d: chunk %zu without a run slot has slot %u", id, c, k.slot);
        }
    }
    if (!p.chunks.empty()) CHECK(p.chunks.back().part + 1 == p.nbParts, "case %d: nbParts %u", id, p.nbParts);
    if (p.chunks.empty()) CHECK(p.nbParts == 0, "case %d: nbParts %u without chunks", id, p.nbParts);
    for (size_t i = 0; i < p.blockSlots.size(); i++) {
        const bool used = p.blockSlots[i].part != kNoSlot;
        CHECK(used ? slotUses[i] >= 2 : (slotUses[i] == 0 && p.blockSlots[i].task == kNoSlot),
              "case %d: slot %zu used by %u chunks", id, i, slotUses[i]);
    }
    if (S == 0) CHECK(p.nbParts == p.chunks.size(), "case %d: no slots merge nothing", id);
    std::vector<uint32_t> partTask(p.nbParts, kNoSlot);
    for (const KmChunk& k : p.chunks)
        if (k.part < p.nbParts) partTask[k.part] = k.task;
    for (uint32_t g = 0; g < nbTasks; g++) {
        CHECK((p.taskPartCnt[g] > 0) == (p.taskChunkCnt[g] > 0), "case %d: task %u part count", id, g);
        if (!p.taskPartCnt[g]) CHECK(p.taskPartBeg[g] == 0, "case %d: task %u part begin", id, g);
        for (uint32_t x = p.taskPartBeg[g]; x < p.taskPartBeg[g] + p.taskPartCnt[g]; x++)
            CHECK(x < p.nbParts && partTask[x] == g, "case %d: task %u part list", id, g);
    }
    uint64_t listed = 0;
    for (uint32_t g = 0; g < nbTasks; g++) listed += p.taskPartCnt[g];
    CHECK(listed == p.nbParts, "case %d: %llu parts listed of %u", id, (unsigned long long)listed, p.nbParts);
}

static void checkCase(const std::vector<uint32_t>& realTiles, const std::vector<uint8_t>& streamed, uint64_t maxWaves,
                      bool passBase, int id, uint32_t wpb = 4, uint32_t spb = 3) {
    const uint32_t n = uint32_t(realTiles.size());
    std::vector<uint32_t> taskTiles(n), base(n);
    uint32_t g0 = 0;
    uint64_t T = 0;
    for (uint32_t g = 0; g < n; g++) {
        base[g] = g0;
        g0 += realTiles[g];
        taskTiles[g] = streamed[g] ? realTiles[g] : 0u;
        T += taskTiles[g];
    }
    if (!passBase)  // prefix-sum numbering: only meaningful when every task streams
        for (uint32_t g = 0; g < n; g++) CHECK(streamed[g], "case %d: test setup", id);
    const KmPartition p = partitionTiles(taskTiles, passBase ? base : std::vector<uint32_t>{}, maxWaves, wpb, spb);
    const uint64_t W = std::min<uint64_t>(std::max<uint64_t>(maxWaves, 1), T);
    CHECK(p.nbWaves == W, "case %d: nbWaves %u expected %llu", id, p.nbWaves, (unsigned long long)W);
    CHECK(p.waveChunkBeg.size() == size_t(p.nbWaves) + 1, "case %d: waveChunkBeg size", id);
    CHECK(p.waveChunkBeg.front() == 0 && p.waveChunkBeg.back() == p.chunks.size(), "case %d: waveChunkBeg ends", id);
    for (uint32_t w = 0; w < p.nbWaves; w++) {
        CHECK(p.waveChunkBeg[w] <= p.waveChunkBeg[w + 1], "case %d: wave %u order", id, w);
        uint64_t tilesOfWave = 0;
        for (uint32_t c = p.waveChunkBeg[w]; c < p.waveChunkBeg[w + 1]; c++)
            tilesOfWave += p.chunks[c].tileEnd - p.chunks[c].tileBeg;
        const uint64_t want = T * (w + 1) / W - T * w / W;
        CHECK(tilesOfWave == want, "case %d: wave %u has %llu tiles, expected %llu", id, w,
              (unsigned long long)tilesOfWave, (unsigned long long)want);
    }
    std::vector<uint32_t> covered(g0, 0u);
    for (size_t c = 0; c < p.chunks.size(); c++) {
        const KmChunk& k = p.chunks[c];
        CHECK(k.task < n, "case %d: chunk %zu task", id, c);
        if (k.task >= n) continue;
        CHECK(streamed[k.task] && taskTiles[k.task] > 0, "case %d: chunk %zu of non-streamed task %u", id, c, k.task);
        CHECK(k.tileBeg < k.tileEnd, "case %d: chunk %zu empty", id, c);
        CHECK(k.tileBeg >= base[k.task] && k.tileEnd <= base[k.task] + realTiles[k.task],
              "case %d: chunk %zu [%u,%u) outside task %u tiles [%u,%u)", id, c, k.tileBeg, k.tileEnd, k.task,
              base[k.task], base[k.task] + realTiles[k.task]);
        for (uint32_t t = k.tileBeg; t < k.tileEnd && t < g0; t++) covered[t]++;
        if (c) CHECK(p.chunks[c - 1].tileEnd <= k.tileBeg, "case %d: chunks %zu not increasing", id, c);
    }
    for (uint32_t g = 0; g < n; g++) {
        for (uint32_t t = base[g]; t < base[g] + realTiles[g]; t++)
            CHECK(covered[t] == (streamed[g] ? 1u : 0u), "case %d: tile %u of task %u covered %u times", id, t, g,
                  covered[t]);
        CHECK((p.taskChunkCnt[g] > 0) == (taskTiles[g] > 0), "case %d: task %u chunk count %u", id, g,
              p.taskChunkCnt[g]);
        for (uint32_t c = p.taskChunkBeg[g]; c < p.taskChunkBeg[g] + p.taskChunkCnt[g]; c++)
            CHECK(c < p.chunks.size() && p.chunks[c].task == g, "case %d: task %u chunk list", id, g);
        if (!p.taskChunkCnt[g]) CHECK(p.taskChunkBeg[g] == 0, "case %d: task %u chunk begin", id, g);
    }
    checkParts(p, n, wpb, spb, id);
}

// ---------------------------------------------------------------------------
// Simulation of one MFMA assign pass's partial-sum bookkeeping (k_assignU8Mfma,
// k_reduceChunks, k_updateU8Part) under a given tile -> wave layout.
//
// A tile contributes (1, h(t)) -- a count and a hashed weight -- to its task.
// Each wave accumulates its tiles of one chunk and, when it leaves the chunk,
// flushes the accumulator: a chunk without a merge slot STORES it into its
// partial (the kernel's plain stores), a chunk with a slot ADDS it into the
// workgroup's LDS slot (atomics), and after the workgroup's barrier every slot
// of an active task is stored into its partial.  Stores of different waves land
// in an order fixed by the simulated timing (last writer wins).  Partials of
// inactive (converged) tasks are never written and keep stale values from
// earlier passes, as on the device.  The reduction then sums each active task's
// partials -- directly for at most kDirectParts, else through reduce items of at
// most kReduceItemParts that must tile the task's partial range -- and the
// result must equal the sum over the task's streamed tiles.
//
// Layouts:
//   OwnRanges: the current kernel; wave w processes its own chunks in order.
//   InTurn:    the reverted variant -- one contiguous tile range per
//              workgroup (the union of its waves' ranges), its waves taking the
//              tiles in turn; which wave takes which tile depends on timing
//              (modelled by a random pick per tile), and so does store order.
// ---------------------------------------------------------------------------
constexpr uint32_t kDirectPartsSim = 32, kReduceItemPartsSim = 32;

struct SimAcc {
    uint64_t cnt = 0, sum = 0;
    SimAcc& operator+=(const SimAcc& o) {
        cnt += o.cnt;
        sum += o.sum;
        return *this;
    }
    bool operator==(const SimAcc& o) const { return cnt == o.cnt && sum == o.sum; }
};
static SimAcc tileValue(uint32_t t) { return SimAcc{1, (uint64_t(t) + 1) * 0x9E3779B97F4A7C15ull}; }

enum class SimLayout { OwnRanges, InTurn };

// Returns the number of active tasks whose reduced totals are wrong after this pass.
static int simulatePass(const KmPartition& p, const std::vector<uint8_t>& active, std::vector<SimAcc>& partials,
                        SimLayout layout, std::mt19937_64& rng, uint32_t* firstBadChunk) {
    const uint32_t W = p.wavesPerBlock, S = p.slotsPerBlock, nbBlocks = (p.nbWaves + W - 1) / W;
    const uint32_t nbTasks = uint32_t(p.taskPartBeg.size());
    struct Store {
        uint64_t time;
        uint32_t part;
        SimAcc v;
    };
    std::vector<Store> stores;
    std::vector<SimAcc> truth(nbTasks);
    for (uint32_t blk = 0; blk < nbBlocks; blk++) {
        const uint32_t w0 = blk * W, w1 = std::min(w0 + W, p.nbWaves);
        std::vector<SimAcc> slot(S);
        // the tiles each wave of this workgroup processes, in order, with their chunk
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> work(W);
        for (uint32_t w = w0; w < w1; w++)
            for (uint32_t c = p.waveChunkBeg[w]; c < p.waveChunkBeg[w + 1]; c++) {
                if (!active[p.chunks[c].task]) continue;  // nextActive: converged tasks' chunks are skipped
                for (uint32_t t = p.chunks[c].tileBeg; t < p.chunks[c].tileEnd; t++) {
                    truth[p.chunks[c].task] += tileValue(t);
                    const uint32_t who = layout == SimLayout::OwnRanges ? w - w0 : uint32_t(rng() % (w1 - w0));
                    work[who].push_back({t, c});
                }
            }
        uint64_t clock = 0;
        for (uint32_t wv = 0; wv < W; wv++) {
            const auto& wk = work[wv];
            uint64_t time = rng() % 64;  // the wave's start
            for (size_t i = 0; i < wk.size();) {
                const uint32_t c = wk[i].second;
                SimAcc acc;
                for (; i < wk.size() && wk[i].second == c; i++) {
                    acc += tileValue(wk[i].first);
                    time += 1 + rng() % 8;
                }
                const KmChunk& k = p.chunks[c];
                if (k.slot == kNoSlot) stores.push_back(Store{time, k.part, acc});
                else slot[k.slot] += acc;
            }
            clock = std::max(clock, time);
        }
        for (uint32_t sl = 0; sl < S; sl++) {  // after the barrier: slots of active tasks leave once
            const KmSlot& ks = p.blockSlots[uint64_t(blk) * S + sl];
            if (ks.part != kNoSlot && active[ks.task]) stores.push_back(Store{clock + 1, ks.part, slot[sl]});
        }
    }
    std::stable_sort(stores.begin(), stores.end(), [](const Store& a, const Store& b) { return a.time < b.time; });
    for (const Store& st : stores) partials[st.part] = st.v;
    int bad = 0;
    for (uint32_t g = 0; g < nbTasks; g++) {
        if (!active[g] || p.taskPartCnt[g] == 0) continue;
        const uint32_t cb = p.taskPartBeg[g], nc = p.taskPartCnt[g];
        SimAcc tot;
        if (nc <= kDirectPartsSim) {
            for (uint32_t x = cb; x < cb + nc; x++) tot += partials[x];
        } else {  // reduce items as the host lists them: they must tile [cb, cb + nc)
            uint32_t next = cb;
            for (uint32_t c = 0; c < nc; c += kReduceItemPartsSim) {
                const uint32_t b = cb + c, e = cb + std::min(nc, c + kReduceItemPartsSim);
                CHECK(b == next && e > b, "reduce items of task %u do not tile its partials", g);
                for (uint32_t x = b; x < e; x++) tot += partials[x];
                next = e;
            }
            CHECK(next == cb + nc, "reduce items of task %u stop early", g);
        }
        if (!(tot == truth[g])) {
            if (bad++ == 0 && firstBadChunk)
                for (uint32_t c = p.taskChunkBeg[g]; c < p.taskChunkBeg[g] + p.taskChunkCnt[g]; c++)
                    if (p.chunks[c].slot == kNoSlot && p.chunks[c].tileEnd - p.chunks[c].tileBeg >= 2) {
                        *firstBadChunk = c;
                        break;
                    }
        }
    }
    return bad;
}

// Several passes over one level with tasks converging at random; returns the
// number of (pass, task) pairs with wrong totals, and the number of passes with any.
static std::pair<uint64_t, uint32_t> simulateLevel(const std::vector<uint32_t>& taskTiles, uint64_t maxWaves,
                                                   uint32_t wpb, uint32_t spb, SimLayout layout, uint32_t passes,
                                                   std::mt19937_64& rng, uint32_t* firstBadChunk = nullptr,
                                                   KmPartition* out = nullptr) {
    const KmPartition p = partitionTiles(taskTiles, {}, maxWaves, wpb, spb);
    std::vector<uint8_t> active(taskTiles.size(), 1);
    std::vector<SimAcc> partials(p.nbParts, SimAcc{0xDEADull, 0xBEEFull});  // never-written partials are garbage
    uint64_t bad = 0;
    uint32_t badPasses = 0;
    for (uint32_t it = 0; it < passes; it++) {
        const int b = simulatePass(p, active, partials, layout, rng, firstBadChunk);
        bad += uint64_t(b);
        badPasses += b > 0;
        for (auto& a : active)
            if (a && rng() % 4 == 0) a = 0;  // converged: its partials go stale, its slots are not written
    }
    if (out) *out = p;
    return {bad, badPasses};
}

static void simulateLayouts(std::mt19937_64& rng) {
    // a hand-sized level: one workgroup of four waves (the kernel's shape), three
    // tasks whose middle one lies inside one wave's range of the per-wave layout
    {
        const std::vector<uint32_t> tiles = {5, 2, 9};  // 16 tiles, 4 per wave: task 1 = tiles 5..6, inside wave 1
        KmPartition p;
        uint32_t badChunk = kNoSlot;
        const auto own = simulateLevel(tiles, 4, 4, 3, SimLayout::OwnRanges, 1, rng, nullptr, &p);
        CHECK(own.first == 0, "per-wave layout loses sums on the hand-sized level");
        uint64_t turnBad = 0;
        for (int r = 0; r < 64; r++) turnBad += simulateLevel(tiles, 4, 4, 3, SimLayout::InTurn, 1, rng, &badChunk).first;
        std::printf("hand-sized level {5,2,9} tiles, 4 waves, 3 slots: chunks");
        for (const KmChunk& c : p.chunks)
            std::printf(" [task %u tiles %u-%u slot %s]", c.task, c.tileBeg, c.tileEnd - 1,
                        c.slot == kNoSlot ? "none" : std::to_string(c.slot).c_str());
        std::printf("\n  per-wave ranges: 0 wrong task totals; in-turn: %llu wrong in 64 runs\n",
                    (unsigned long long)turnBad);
        CHECK(turnBad > 0, "the in-turn layout should lose the sums of task 1 (a chunk without a slot)");
    }
    // random levels with the kernel's shape (4 waves, 3 slots), several passes each
    uint64_t ownBad = 0, turnBad = 0;
    uint32_t levels = 0, turnLevelsBad = 0;
    for (int r = 0; r < 300; r++) {
        const uint32_t n = 1 + uint32_t(rng() % (r < 150 ? 60 : 3000));
        std::vector<uint32_t> tiles(n);
        for (uint32_t g = 0; g < n; g++) {
            const uint32_t pick = uint32_t(rng() % 10);
            tiles[g] = pick == 0 ? 0u : pick < 7 ? 1u + uint32_t(rng() % 24) : 17u + uint32_t(rng() % 400);
        }
        const uint64_t mw = std::vector<uint64_t>{4, 8, 64, 512, 2048}[rng() % 5];
        ownBad += simulateLevel(tiles, mw, 4, 3, SimLayout::OwnRanges, 6, rng).first;
        const auto t = simulateLevel(tiles, mw, 4, 3, SimLayout::InTurn, 6, rng);
        turnBad += t.first;
        turnLevelsBad += t.first > 0;
        levels++;
    }
    std::printf("random levels (4 waves, 3 slots, 6 passes each): per-wave ranges %llu wrong task totals in %u "
                "levels; in-turn %llu wrong in %u of %u levels\n",
                (unsigned long long)ownBad, levels, (unsigned long long)turnBad, turnLevelsBad, levels);
    CHECK(ownBad == 0, "the current layout lost sums");
    CHECK(turnBad > 0, "the simulation does not detect the in-turn layout's lost stores");
}

int main() {
    std::mt19937_64 rng(2024);
    int cases = 0;
    // hand-picked shapes: leading, inner and trailing non-streamed tasks, empty tasks
    const std::vector<std::vector<uint32_t>> shapes = {{5}, {1, 1, 1}, {0, 3, 0, 2, 0}, {40, 2, 2, 40}, {7, 0, 0, 9}};
    const std::vector<std::vector<uint8_t>> masks = {{1}, {1, 0, 1}, {1, 1, 1, 1, 1}, {1, 0, 0, 1}, {0, 1, 1, 1}};
    for (size_t i = 0; i < shapes.size(); i++)
        for (uint64_t mw : {1ull, 2ull, 3ull, 5ull, 64ull, 4096ull})
            for (uint32_t wpb : {1u, 2u, 4u, 8u})
                for (uint32_t spb : {0u, 1u, 2u, 7u}) {
                    checkCase(shapes[i], masks[i], mw, true, cases++, wpb, spb);
                    std::vector<uint8_t> all(shapes[i].size(), 1);
                    checkCase(shapes[i], all, mw, false, cases++, wpb, spb);
                }
    // random levels: sizes from a heavy tail, about half the tasks not streamed
    for (int r = 0; r < 400; r++) {
        const uint32_t n = 1 + uint32_t(rng() % (r < 200 ? 40 : 4096));
        std::vector<uint32_t> tiles(n);
        std::vector<uint8_t> st(n);
        for (uint32_t g = 0; g < n; g++) {
            const uint32_t pick = uint32_t(rng() % 10);
            tiles[g] = pick == 0 ? 0u : pick < 6 ? 1u + uint32_t(rng() % 32) : 1u + uint32_t(rng() % 3000);
            st[g] = (tiles[g] > 32 || rng() % 2) ? 1 : 0;
        }
        const uint64_t mw = std::vector<uint64_t>{1, 5, 256, 1024, 8192}[rng() % 5];
        const uint32_t wpb = std::vector<uint32_t>{1, 2, 4, 4, 8}[rng() % 5];
        const uint32_t spb = std::vector<uint32_t>{0, 1, 2, 2, 3, 7}[rng() % 6];
        checkCase(tiles, st, mw, true, cases++, wpb, spb);
        std::vector<uint8_t> all(n, 1);
        checkCase(tiles, all, mw, rng() % 2 == 0, cases++, wpb, spb);
    }
    simulateLayouts(rng);
    std::printf("%d cases, %d failures\n", cases, fails);
    return fails ? EXIT_FAILURE : EXIT_SUCCESS;
}
