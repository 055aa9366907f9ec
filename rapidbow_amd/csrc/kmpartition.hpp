// kmpartition.hpp -- host-only work partition of the MFMA k-means assign
// (no HIP dependency, so it is also compiled by plain g++ for its tests).
//
// The level's tiles are numbered task after task.  Only the streamed tiles
// are partitioned: taskTiles[g] may be smaller than the task's real tile count
// (0 for a task whose Lloyd loop runs elsewhere), and tileBeg[g] is the global
// index of task g's first tile.  The streamed tiles are cut into nbWaves equal
// ranges, one per resident wave, and each range at task boundaries into
// chunks.  Wave w processes chunks [waveChunkBeg[w], waveChunkBeg[w+1]); a
// task's chunks are contiguous; chunk tile indices are global, so they address
// the level's tile and operand arrays.
//
// Partial sums.  Waves are launched wavesPerBlock to a workgroup.  Within one
// workgroup, consecutive chunks of the same task (they necessarily belong to
// different waves: the last chunk of one wave and the first of the next) are
// merged in LDS and leave as ONE partial; every other chunk writes its own.
// A workgroup has at most wavesPerBlock - 1 such runs (each crosses a wave
// boundary of its own); the first slotsPerBlock of them get an LDS slot and
// any further one keeps a partial per chunk.  Partials are numbered in chunk
// order, so the partials of a task are contiguous
// [taskPartBeg[g], taskPartBeg[g] + taskPartCnt[g]).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace rbow_amd {

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

struct KmChunk {       // a run of tiles of one task, processed by one wave
    uint32_t task;
    uint32_t tileBeg;  // global tile indices [tileBeg, tileEnd)
    uint32_t tileEnd;
    uint32_t part;     // partial-sum index this chunk contributes to
    uint32_t slot;     // LDS merge slot of its workgroup, or kNoSlot (writes partial `part` itself)
};

struct KmSlot {        // a workgroup's merge slot: the partial it becomes and its task
    uint32_t part;     // kNoSlot: unused
    uint32_t task;
};

struct KmPartRange {   // a reduce item: partials [partBeg, partEnd) of one task
    uint32_t task;
    uint32_t partBeg;
    uint32_t partEnd;
};

struct KmPartition {
    uint32_t nbWaves = 0, wavesPerBlock = 1, slotsPerBlock = 0, nbParts = 0;
    std::vector<KmChunk> chunks;
    std::vector<uint32_t> waveChunkBeg;               // [nbWaves + 1]
    std::vector<uint32_t> taskChunkBeg, taskChunkCnt;  // [task]; 0, 0 for tasks with no streamed tile
    std::vector<uint32_t> taskPartBeg, taskPartCnt;    // [task]; 0, 0 for tasks with no streamed tile
    std::vector<KmSlot> blockSlots;                    // [block][slotsPerBlock]
};

// taskTiles[g]: streamed tiles of task g (its first taskTiles[g] tiles);
// tileBeg[g]: global first tile of task g (empty: the prefix sums of taskTiles);
// maxWaves: resident waves of the assign kernel; wavesPerBlock: its waves per
// workgroup; slotsPerBlock: its LDS merge slots (at most wavesPerBlock - 1 are used).
inline KmPartition partitionTiles(const std::vector<uint32_t>& taskTiles, const std::vector<uint32_t>& tileBeg,
                                  uint64_t maxWaves, uint32_t wavesPerBlock = 1, uint32_t slotsPerBlock = 0) {
    KmPartition p;
    const uint32_t nbTasks = uint32_t(taskTiles.size());
    std::vector<uint32_t> base(tileBeg);
    if (base.empty()) {
        base.resize(nbTasks);
        uint32_t t = 0;
        for (uint32_t g = 0; g < nbTasks; g++) {
            base[g] = t;
            t += taskTiles[g];
        }
    }
    uint64_t T = 0;
    for (uint32_t n : taskTiles) T += n;
    p.taskChunkBeg.assign(nbTasks, 0u);
    p.taskChunkCnt.assign(nbTasks, 0u);
    p.nbWaves = uint32_t(std::min<uint64_t>(std::max<uint64_t>(maxWaves, 1), T));
    p.waveChunkBeg.assign(p.nbWaves + 1, 0u);
    uint32_t g = 0;     // task holding streamed position lo
    uint64_t gBeg = 0;  // its first streamed position
    for (uint32_t w = 0; w < p.nbWaves; w++) {
        p.waveChunkBeg[w] = uint32_t(p.chunks.size());
        uint64_t lo = T * w / p.nbWaves;
        const uint64_t hi = T * (w + 1) / p.nbWaves;
        while (lo < hi) {
            while (lo >= gBeg + taskTiles[g]) gBeg += taskTiles[g++];  // lo < T: some later task holds it
            const uint64_t end = std::min<uint64_t>(hi, gBeg + taskTiles[g]);
            if (p.taskChunkCnt[g] == 0) p.taskChunkBeg[g] = uint32_t(p.chunks.size());
            p.chunks.push_back(KmChunk{g, uint32_t(base[g] + (lo - gBeg)), uint32_t(base[g] + (end - gBeg)), 0u, kNoSlot});
            p.taskChunkCnt[g]++;
            lo = end;
        }
    }
    p.waveChunkBeg[p.nbWaves] = uint32_t(p.chunks.size());

    // partials: runs of same-task chunks inside one workgroup share one
    p.wavesPerBlock = std::max(wavesPerBlock, 1u);
    p.slotsPerBlock = std::min(slotsPerBlock, p.wavesPerBlock - 1);
    const uint32_t W = p.wavesPerBlock, S = p.slotsPerBlock;
    const uint32_t nbBlocks = (p.nbWaves + W - 1) / W;
    p.blockSlots.assign(uint64_t(nbBlocks) * S, KmSlot{kNoSlot, kNoSlot});
    p.taskPartBeg.assign(nbTasks, 0u);
    p.taskPartCnt.assign(nbTasks, 0u);
    uint32_t part = 0;
    for (uint32_t blk = 0; blk < nbBlocks; blk++) {
        const uint32_t cb = p.waveChunkBeg[blk * W], ce = p.waveChunkBeg[std::min((blk + 1) * W, p.nbWaves)];
        uint32_t nextSlot = 0;
        for (uint32_t c = cb; c < ce;) {
            const uint32_t task = p.chunks[c].task;
            uint32_t e = c + 1;
            while (e < ce && p.chunks[e].task == task) e++;
            if (p.taskPartCnt[task] == 0) p.taskPartBeg[task] = part;
            // runs beyond the workgroup's slots keep a partial per chunk
            if (e - c >= 2 && nextSlot < S) {
                const uint32_t slot = nextSlot++;
                p.blockSlots[uint64_t(blk) * S + slot] = KmSlot{part, task};
                for (uint32_t x = c; x < e; x++) {
                    p.chunks[x].part = part;
                    p.chunks[x].slot = slot;
                }
                p.taskPartCnt[task]++;
                part++;
            } else {
                for (uint32_t x = c; x < e; x++) {
                    p.chunks[x].part = part++;
                    p.taskPartCnt[task]++;
                }
            }
            c = e;
        }
    }
    p.nbParts = part;
    return p;
}

}  // namespace rbow_amd
