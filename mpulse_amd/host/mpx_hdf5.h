// HDF5 output in the on-disk form of MPulse's HDFostream (src/hdfstream.{h,t,cpp}), written
// without the HDF5 library (none is installed with this build).
//
// The files have the structure the HDF5 1.6 library gives such a stream with its default
// properties, taken from the published HDF5 file-format specification and checked byte for
// byte against the file the reference's own test program wrote (test/test.h5): a version-0
// superblock (offsets and lengths of 8 bytes, group K 4 / 16), the root group as a symbol
// table (local heap of names, one version-1 group B-tree node, symbol-table nodes of up to 8
// entries), and per dataset a version-1 object header with a fill-value, an IEEE float64
// little-endian datatype, a simple dataspace, a contiguous layout and a modification-time
// message, optionally an int32 "extent" attribute (the older stream's low/high indices).
// Metadata comes first, in 2 KiB blocks; the raw data of each dataset follows in write order.
//
// Dataset names follow HDFstream::getNextBlockName (src/hdfstream.cpp): "data0", "data1", ...,
// or after setBlockName(n): "n", "n1", "n2", ...
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace mpulse_hip {

class HDFostream {
  public:
    // capacity: datasets the file can take (their metadata is reserved when the file opens)
    HDFostream() = default;
    ~HDFostream();
    HDFostream(const HDFostream &) = delete;
    HDFostream &operator=(const HDFostream &) = delete;

    bool open(const std::string &path, int capacity = 64);
    // One dataset of rank 1..3 with HDF5 dimensions `dims` (slowest first), float64 in C order
    // over `dims`; `extent`, when not empty, becomes the int32 "extent" attribute.  The file
    // is a complete HDF5 file again when the call returns.
    bool write(const std::vector<uint64_t> &dims, const double *data, const std::vector<int32_t> &extent = {});
    void setBlockName(const std::string &name);  // HDFstream::setBlockName
    void setModificationTime(uint32_t t) { mtime_ = t; }  // 0: the time of each write
    void close();
    bool good() const { return f_ != nullptr; }
    const std::string &error() const { return err_; }

  private:
    struct Set {
        std::string name;
        uint32_t heapOffset;
        uint64_t header, raw;
        std::vector<uint64_t> dims;
        std::vector<int32_t> extent;
        uint32_t mtime;
    };
    std::string nextBlockName();
    bool writeMetadata();
    bool fail(const std::string &m);

    FILE *f_ = nullptr;
    std::string err_;
    std::string blockname_ = "data";
    int setsCount_ = 0;
    int capacity_ = 0;
    uint32_t mtime_ = 0;
    uint64_t heapData_ = 0, heapSize_ = 0, btree_ = 0, rootHeader_ = 0, rawStart_ = 0, eof_ = 0;
    uint32_t heapUsed_ = 8;  // offset 0 holds the empty name
    std::vector<uint64_t> headerSlot_, snodSlot_;
    std::vector<Set> sets_;
};

}  // namespace mpulse_hip
