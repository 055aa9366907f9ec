// kmpartition.hpp -- host-only work partition of the MFMA k-means assign
// (no HIP dependency, so it is also compiled by plain g++ for its tests).
//
// The level's tiles are numbered task after task.  Only the streamed tiles
// are partitioned: taskTiles[g] may be smaller than the task's real tile count
// (0 for a task whose Lloyd loop runs elsewhere), and tileBeg[g] is the global
// index of task g's first tile.  The streamed tiles are cut into nbWaves equal
// ranges, one per resident wave, and each range at task boundaries into
// chunks (one partial-sum slot each).  Wave w processes chunks
// [waveChunkBeg[w], waveChunkBeg[w+1]); a task's chunks are contiguous; chunk
// tile indices are global, so they address the level's tile and operand arrays.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace rbow_amd {

struct KmChunk {       // a run of tiles of one task, processed by one wave
    uint32_t task;
    uint32_t tileBeg;  // global tile indices [tileBeg, tileEnd)
    uint32_t tileEnd;
};

struct KmPartition {
    uint32_t nbWaves = 0;
    std::vector<KmChunk> chunks;
    std::vector<uint32_t> waveChunkBeg;               // [nbWaves + 1]
    std::vector<uint32_t> taskChunkBeg, taskChunkCnt;  // [task]; 0, 0 for tasks with no streamed tile
};

// taskTiles[g]: streamed tiles of task g (its first taskTiles[g] tiles);
// tileBeg[g]: global first tile of task g (empty: the prefix sums of taskTiles);
// maxWaves: resident waves of the assign kernel.
inline KmPartition partitionTiles(const std::vector<uint32_t>& taskTiles, const std::vector<uint32_t>& tileBeg,
                                  uint64_t maxWaves) {
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
            p.chunks.push_back(KmChunk{g, uint32_t(base[g] + (lo - gBeg)), uint32_t(base[g] + (end - gBeg))});
            p.taskChunkCnt[g]++;
            lo = end;
        }
    }
    p.waveChunkBeg[p.nbWaves] = uint32_t(p.chunks.size());
    return p;
}

}  // namespace rbow_amd
