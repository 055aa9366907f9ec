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
// each task's partials.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
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
    std::printf("%d cases, %d failures\n", cases, fails);
    return fails ? EXIT_FAILURE : EXIT_SUCCESS;
}
