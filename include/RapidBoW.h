/*
 * RapidBoW.h -- header-only C++ facade over the C ABI (rapidbow.h).
 *
 * Recreates the reference's public C++ surface (RapidBoW.h:29-109 of
 * lowsfer/RapidBoW) -- names, ownership (std::unique_ptr) and exception
 * types/messages -- so callers of the reference compile against this header.
 * The one signature change is the stream type: hipStream_t replaces
 * cudaStream_t (the library runs on HIP / gfx950).
 */
#ifndef RAPIDBOW_AMD_RAPIDBOW_HPP
#define RAPIDBOW_AMD_RAPIDBOW_HPP

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rapidbow.h"

namespace rbow {

enum class DataType { kBinary, kUInt8 };             // RapidBoW.h:31-34
enum class DistanceType { kHamming, kL1, kL2 };      // RapidBoW.h:36-41

struct DescriptorAttributes {                        // RapidBoW.h:43-48
    DataType dataType;
    DistanceType distanceType;
    size_t nbDims;
};

constexpr DescriptorAttributes SiftAttr{DataType::kUInt8, DistanceType::kL2, 128u};  // RapidBoW.h:50

constexpr inline size_t getDescriptorBytes(DescriptorAttributes descAttr) {  // RapidBoW.h:52-58
    return descAttr.dataType == DataType::kBinary ? (descAttr.nbDims + 7) / 8 : descAttr.nbDims;
}

namespace detail {
inline rbow_descriptor_attributes toC(DescriptorAttributes a) {
    return rbow_descriptor_attributes{a.dataType == DataType::kUInt8 ? RBOW_DATA_UINT8 : RBOW_DATA_BINARY,
                                      a.distanceType == DistanceType::kL2   ? RBOW_DIST_L2
                                      : a.distanceType == DistanceType::kL1 ? RBOW_DIST_L1
                                                                            : RBOW_DIST_HAMMING,
                                      uint64_t(a.nbDims)};
}
// Reference errors are std::runtime_error with fixed messages; the C ABI
// carries them as status + message.
inline void check(rbow_status s) {
    if (s != RBOW_OK) throw std::runtime_error(rbow_last_error());
}
inline rbow_stream toC(hipStream_t s) { return reinterpret_cast<rbow_stream>(s); }
}  // namespace detail

class IVocabulary {  // RapidBoW.h:60-71
public:
    virtual DescriptorAttributes getDescriptorAttributes() const = 0;
    virtual uint32_t getBranchFactor() const = 0;
    virtual uint32_t getNbLevels() const = 0;
    // descriptors and indicesInLeafLevel must be device memory
    virtual void lookUp(const void* descriptors, uint32_t nbDesc, uint32_t* indicesInLeafLevel,
                        hipStream_t stream) const = 0;
    virtual std::vector<std::uint8_t> serialize() const = 0;
    virtual ~IVocabulary() = default;
};

// Concrete vocabulary owning a C handle.
class Vocabulary final : public IVocabulary {
public:
    explicit Vocabulary(rbow_vocabulary* h) : mH(h) {}
    ~Vocabulary() override { rbow_voc_destroy(mH); }
    Vocabulary(const Vocabulary&) = delete;
    Vocabulary& operator=(const Vocabulary&) = delete;
    DescriptorAttributes getDescriptorAttributes() const override { return SiftAttr; }
    uint32_t getBranchFactor() const override { return rbow_voc_branch_factor(mH); }
    uint32_t getNbLevels() const override { return rbow_voc_nb_levels(mH); }
    void lookUp(const void* descriptors, uint32_t nbDesc, uint32_t* indicesInLeafLevel,
                hipStream_t stream) const override {
        detail::check(rbow_voc_lookup(mH, descriptors, nbDesc, indicesInLeafLevel, detail::toC(stream)));
    }
    std::vector<std::uint8_t> serialize() const override {
        size_t n = 0;
        detail::check(rbow_voc_serialize(mH, nullptr, 0, &n));
        std::vector<std::uint8_t> b(n);
        detail::check(rbow_voc_serialize(mH, b.data(), b.size(), &n));
        return b;
    }
    // RapidBoWInternal.h:29-30 getLevelWeights
    const void* getLevelWeights(uint32_t idxLevel) const { return rbow_voc_level_weights(mH, idxLevel); }
    rbow_vocabulary* handle() const { return mH; }

private:
    rbow_vocabulary* mH;
};

namespace impl {
// RapidBoW.h:75-80; unsupported attributes -> nullptr (RapidBoW.cpp:37-40)
inline IVocabulary* buildVocabularyImpl(DescriptorAttributes descAttr, uint32_t branchFactor, const void* devDesc,
                                        uint32_t nbDesc, uint32_t nbDoc, uint32_t nbLevels, hipStream_t stream) {
    rbow_vocabulary* h = nullptr;
    const rbow_status s = rbow_build_vocabulary(detail::toC(descAttr), branchFactor, devDesc, nbDesc, nbDoc,
                                                nbLevels, detail::toC(stream), &h);
    if (s == RBOW_ERR_UNSUPPORTED) return nullptr;
    detail::check(s);
    return new Vocabulary(h);
}
}  // namespace impl

inline std::unique_ptr<IVocabulary> buildVocabulary(DescriptorAttributes descAttr, uint32_t branchFactor,
                                                    const void* devDesc, uint32_t nbDesc, uint32_t nbDoc,
                                                    uint32_t nbLevels, hipStream_t stream) {  // RapidBoW.h:82-88
    return std::unique_ptr<IVocabulary>{
        impl::buildVocabularyImpl(descAttr, branchFactor, devDesc, nbDesc, nbDoc, nbLevels, stream)};
}

inline std::unique_ptr<IVocabulary> deserializeVocabulary(const std::uint8_t* blob, size_t size) {  // RapidBoW.h:90
    rbow_vocabulary* h = nullptr;
    detail::check(rbow_voc_deserialize(blob, size, &h));
    return std::make_unique<Vocabulary>(h);
}

class IDataBase {  // RapidBoW.h:92-105
public:
    virtual const IVocabulary* getVocabulary() const = 0;
    // hostIndicesInLeafLevel is host memory, unlike in IVocabulary::lookUp
    virtual void addDoc(uint32_t idxDoc, const uint32_t* hostIndicesInLeafLevel, uint32_t nbDesc) = 0;
    virtual std::vector<uint32_t> query(const uint32_t* hostIndicesInLeafLevel, uint32_t nbDesc,
                                        uint32_t maxNbResults) const = 0;
    virtual std::vector<uint32_t> queryAndAddDoc(uint32_t idxDoc, const uint32_t* hostIndicesInLeafLevel,
                                                 uint32_t nbDesc, uint32_t maxNbResults) = 0;
    virtual ~IDataBase() = default;
};

class DataBase final : public IDataBase {
public:
    // DataBase.h:43-45: borrows the vocabulary, which must outlive the database;
    // a foreign IVocabulary implementation throws std::bad_cast.
    explicit DataBase(const IVocabulary& voc) : mVoc(dynamic_cast<const Vocabulary&>(voc)) {
        detail::check(rbow_db_create(mVoc.handle(), &mH));
    }
    ~DataBase() override { rbow_db_destroy(mH); }
    DataBase(const DataBase&) = delete;
    DataBase& operator=(const DataBase&) = delete;
    const IVocabulary* getVocabulary() const override { return &mVoc; }
    void addDoc(uint32_t idxDoc, const uint32_t* leaf, uint32_t nbDesc) override {
        detail::check(rbow_db_add_doc(mH, idxDoc, leaf, nbDesc));
    }
    std::vector<uint32_t> query(const uint32_t* leaf, uint32_t nbDesc, uint32_t maxNbResults) const override {
        std::vector<uint32_t> r(maxNbResults);
        uint32_t n = 0;
        detail::check(rbow_db_query(mH, leaf, nbDesc, maxNbResults, r.data(), nullptr, &n));
        r.resize(n);
        return r;
    }
    std::vector<uint32_t> queryAndAddDoc(uint32_t idxDoc, const uint32_t* leaf, uint32_t nbDesc,
                                         uint32_t maxNbResults) override {
        std::vector<uint32_t> r(maxNbResults);
        uint32_t n = 0;
        detail::check(rbow_db_query_and_add_doc(mH, idxDoc, leaf, nbDesc, maxNbResults, r.data(), &n));
        r.resize(n);
        return r;
    }
    rbow_database* handle() const { return mH; }

private:
    const Vocabulary& mVoc;
    rbow_database* mH = nullptr;
};

inline std::unique_ptr<IDataBase> createDataBase(const IVocabulary* voc) {  // RapidBoW.h:107
    return std::make_unique<DataBase>(*voc);
}

}  // namespace rbow

#endif
