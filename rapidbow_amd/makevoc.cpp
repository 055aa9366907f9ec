// makevoc -- builds a vocabulary from a raw file of SIFT descriptors and writes
// its blob: the command-line tool of the reference (main.cpp:25-49), written
// against this library's C++ facade (include/RapidBoW.h).
//
//   makevoc <desc_file> <voc_file> [--levels L] [--docs N]
//
// desc_file holds uint8[128] descriptors back to back.  The defaults are the
// reference's fixed build: branch 16, L = 5, nbDoc = 527 (main.cpp:25-26), with
// the reference's seeding and pass cap.  The blob is the reference format
// (Vocabulary.cpp:175-197).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "RapidBoW.h"

namespace {

int usage() {
    std::fprintf(stderr, "usage: makevoc <desc_file> <voc_file> [--levels L] [--docs N]\n");
    return EXIT_FAILURE;
}

std::vector<uint8_t> readAll(const char* path) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    const std::streamsize n = in.tellg();
    std::vector<uint8_t> data(size_t(n > 0 ? n : 0));
    in.seekg(0);
    if (n > 0 && !in.read(reinterpret_cast<char*>(data.data()), n))
        throw std::runtime_error(std::string("cannot read ") + path);
    return data;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return usage();
    uint32_t levels = 5, docs = 527;
    for (int i = 3; i < argc; i++) {
        if (i + 1 < argc && !std::strcmp(argv[i], "--levels")) levels = uint32_t(std::strtoul(argv[++i], nullptr, 10));
        else if (i + 1 < argc && !std::strcmp(argv[i], "--docs")) docs = uint32_t(std::strtoul(argv[++i], nullptr, 10));
        else return usage();
    }
    try {
        const std::vector<uint8_t> desc = readAll(argv[1]);
        if (desc.size() % 128 != 0) throw std::runtime_error("descriptor file size is not a multiple of 128 bytes");
        const uint32_t nbDesc = uint32_t(desc.size() / 128);
        void* dev = nullptr;
        if (hipMalloc(&dev, desc.size() ? desc.size() : 1) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        struct Free {
            void* p;
            ~Free() { (void)hipFree(p); }
        } guard{dev};
        if (!desc.empty() && hipMemcpy(dev, desc.data(), desc.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        const auto voc = rbow::buildVocabulary(rbow::SiftAttr, 16, dev, nbDesc, docs, levels, nullptr);
        if (!voc) throw std::runtime_error("unsupported descriptor attributes");
        const std::vector<uint8_t> blob = voc->serialize();
        std::ofstream out(argv[2], std::ios::binary | std::ios::trunc);
        if (!out.write(reinterpret_cast<const char*>(blob.data()), std::streamsize(blob.size())))
            throw std::runtime_error(std::string("cannot write ") + argv[2]);
        std::printf("makevoc: %u descriptors, %u levels -> %zu bytes\n", nbDesc, levels, blob.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "makevoc: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
