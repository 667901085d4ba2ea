// HDF5 output in the on-disk form of MPulse's HDFostream, without the HDF5 library.
// Structures and field layouts are those of the HDF5 file-format specification (version 0
// superblock, local heap, version-1 B-tree and symbol-table nodes, version-1 object headers).
#include "mpx_hdf5.h"

#include <algorithm>
#include <cstring>
#include <ctime>

namespace mpulse_hip {

namespace {

constexpr uint64_t UNDEF = ~0ull;
constexpr int LEAF_K = 4, INTERNAL_K = 16;  // group B-tree K values of the library's defaults
constexpr uint64_t SUPER = 96, HEAP_HDR = 32, BTREE = 24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8;
constexpr uint64_t ROOT_HDR = 16 + 32, DSET_HDR = 16 + 256, SNOD = 8 + 2 * LEAF_K * 40, META_BLOCK = 2048;

uint64_t pad8(uint64_t n) { return (n + 7) & ~7ull; }

struct Image {
    std::vector<unsigned char> b;
    void put(uint64_t at, uint64_t v, int n)
    {
        for (int i = 0; i < n; ++i) b[at + i] = (unsigned char)(v >> (8 * i));
    }
    void bytes(uint64_t at, const void *p, size_t n) { std::memcpy(&b[at], p, n); }
};

// message header: type, body size, flags
uint64_t msg(Image &im, uint64_t at, int type, uint64_t size, int flags)
{
    im.put(at, type, 2);
    im.put(at + 2, size, 2);
    im.put(at + 4, flags, 1);
    return at + 8;
}

}  // namespace

HDFostream::~HDFostream() { close(); }

bool HDFostream::fail(const std::string &m)
{
    err_ = m;
    return false;
}

bool HDFostream::open(const std::string &path, int capacity)
{
    close();
    if (capacity < 1 || capacity > 2 * LEAF_K * 2 * INTERNAL_K) return fail("capacity must be 1..256 datasets");
    f_ = std::fopen(path.c_str(), "wb+");
    if (!f_) return fail("cannot open " + path);
    capacity_ = capacity;
    sets_.clear();
    blockname_ = "data";
    setsCount_ = 0;
    heapUsed_ = 8;
    // the library's default local heap of a new group is 256 bytes; more names get more room
    heapSize_ = std::max<uint64_t>(256, pad8(8 + 16 * (uint64_t)capacity));
    heapData_ = SUPER + HEAP_HDR;
    btree_ = heapData_ + heapSize_;
    rootHeader_ = btree_ + BTREE;
    uint64_t at = rootHeader_ + ROOT_HDR;
    headerSlot_.assign(capacity, 0);
    snodSlot_.clear();
    for (int d = 0; d < capacity; ++d) {
        headerSlot_[d] = at;
        at += DSET_HDR;
        if (d % (2 * LEAF_K) == 0) {  // a symbol-table node follows the header of its first entry
            snodSlot_.push_back(at);
            at += SNOD;
        }
    }
    rawStart_ = (at + META_BLOCK - 1) / META_BLOCK * META_BLOCK;
    eof_ = rawStart_;
    if (!writeMetadata()) return false;
    return true;
}

void HDFostream::setBlockName(const std::string &name)
{
    blockname_ = name;
    setsCount_ = -1;
}

std::string HDFostream::nextBlockName()
{
    // HDFstream::getNextBlockName, src/hdfstream.cpp
    std::string n = blockname_;
    if (setsCount_ < 0)
        setsCount_ = 1;
    else
        n += std::to_string(setsCount_++);
    return n;
}

bool HDFostream::write(const std::vector<uint64_t> &dims, const double *data, const std::vector<int32_t> &extent)
{
    if (!f_) return fail("stream not open");
    if ((int)sets_.size() >= capacity_) return fail("the file holds no more datasets (capacity reached)");
    if (dims.empty() || dims.size() > 3) return fail("dataset rank must be 1..3");
    if (extent.size() > 16) return fail("extent attribute too long");
    Set s;
    s.name = nextBlockName();
    uint64_t nlen = pad8(s.name.size() + 1);
    if (heapUsed_ + nlen > heapSize_) return fail("dataset names exceed the group's name heap");
    uint64_t count = 1;
    for (uint64_t d : dims) count *= d;
    s.heapOffset = heapUsed_;
    s.header = headerSlot_[sets_.size()];
    s.raw = eof_;
    s.dims = dims;
    s.extent = extent;
    s.mtime = mtime_ ? mtime_ : (uint32_t)std::time(nullptr);
    if (std::fseek(f_, (long)s.raw, SEEK_SET) != 0) return fail("seek failed");
    if (count && std::fwrite(data, sizeof(double), count, f_) != count) return fail("write of the raw data failed");
    heapUsed_ += (uint32_t)nlen;
    eof_ += 8 * count;
    sets_.push_back(s);
    return writeMetadata();
}

bool HDFostream::writeMetadata()
{
    Image im;
    im.b.assign(rawStart_, 0);
    // superblock, version 0
    static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    im.bytes(0, sig, 8);
    im.put(13, 8, 1);  // size of offsets
    im.put(14, 8, 1);  // size of lengths
    im.put(16, LEAF_K, 2);
    im.put(18, INTERNAL_K, 2);
    im.put(20, 3, 4);  // file consistency flags, as the library leaves them in a closed file
    im.put(24, 0, 8);  // base address
    im.put(32, UNDEF, 8);
    im.put(40, eof_, 8);
    im.put(48, UNDEF, 8);
    // root group symbol-table entry: name offset 0, header, cache type 1 (B-tree, heap)
    im.put(64, rootHeader_, 8);
    im.put(72, 1, 4);
    im.put(80, btree_, 8);
    im.put(88, SUPER, 8);
    // local heap
    im.bytes(SUPER, "HEAP", 4);
    im.put(SUPER + 8, heapSize_, 8);
    if (heapSize_ - heapUsed_ >= 16) {
        im.put(SUPER + 16, heapUsed_, 8);
        im.put(heapData_ + heapUsed_, 1, 8);  // last free block
        im.put(heapData_ + heapUsed_ + 8, heapSize_ - heapUsed_, 8);
    } else {
        im.put(SUPER + 16, UNDEF, 8);
    }
    im.put(SUPER + 24, heapData_, 8);
    for (const Set &s : sets_) im.bytes(heapData_ + s.heapOffset, s.name.c_str(), s.name.size());
    // symbol-table nodes: the names in strcmp order, 2K per node; the group B-tree above them
    std::vector<const Set *> sorted;
    for (const Set &s : sets_) sorted.push_back(&s);
    std::sort(sorted.begin(), sorted.end(), [](const Set *a, const Set *b) { return a->name < b->name; });
    const size_t per = 2 * LEAF_K, nodes = (sorted.size() + per - 1) / per;
    im.bytes(btree_, "TREE", 4);
    im.put(btree_ + 6, nodes, 2);
    im.put(btree_ + 8, UNDEF, 8);
    im.put(btree_ + 16, UNDEF, 8);
    uint64_t p = btree_ + 24;
    im.put(p, 0, 8);  // key 0: the empty name
    for (size_t j = 0; j < nodes; ++j) {
        uint64_t sn = snodSlot_[j];
        size_t lo = j * per, hi = std::min(sorted.size(), lo + per);
        im.bytes(sn, "SNOD", 4);
        im.put(sn + 4, 1, 1);
        im.put(sn + 6, hi - lo, 2);
        for (size_t e = lo; e < hi; ++e) {
            uint64_t q = sn + 8 + 40 * (e - lo);
            im.put(q, sorted[e]->heapOffset, 8);
            im.put(q + 8, sorted[e]->header, 8);
        }
        im.put(p + 8, sn, 8);
        im.put(p + 16, sorted[hi - 1]->heapOffset, 8);  // key: the largest name of the child
        p += 16;
    }
    // root group object header: symbol-table message and a null message
    im.put(rootHeader_, 1, 1);
    im.put(rootHeader_ + 2, 2, 2);
    im.put(rootHeader_ + 4, 1, 4);
    im.put(rootHeader_ + 8, 32, 4);
    uint64_t m = msg(im, rootHeader_ + 16, 0x11, 16, 1);
    im.put(m, btree_, 8);
    im.put(m + 8, SUPER, 8);
    // dataset object headers
    for (const Set &s : sets_) {
        const int rank = (int)s.dims.size();
        uint64_t h = s.header, q = h + 16, end = h + DSET_HDR;
        int nmsg = 0;
        // fill value (version 1: late allocation, fill written if set, default fill)
        q = msg(im, q, 0x05, 8, 1);
        const unsigned char fill[8] = {1, 2, 2, 1, 0, 0, 0, 0};
        im.bytes(q, fill, 8);
        q += 8, ++nmsg;
        // datatype: IEEE float64 little-endian
        q = msg(im, q, 0x03, 24, 1);
        const unsigned char f64[20] = {0x11, 0x20, 0x3f, 0x00, 8, 0, 0, 0, 0, 0, 64, 0, 52, 11, 0, 52, 0xff, 0x03, 0, 0};
        im.bytes(q, f64, 20);
        q += 24, ++nmsg;
        // dataspace, version 1, no maximum dimensions
        q = msg(im, q, 0x01, 8 + 8 * rank, 0);
        im.put(q, 1, 1);
        im.put(q + 1, rank, 1);
        for (int d = 0; d < rank; ++d) im.put(q + 8 + 8 * d, s.dims[d], 8);
        q += 8 + 8 * rank, ++nmsg;
        // layout, version 2, contiguous: address, dimensions and the element size
        uint64_t lsize = pad8(16 + 4 * (rank + 1));
        q = msg(im, q, 0x08, lsize, 1);
        im.put(q, 2, 1);
        im.put(q + 1, rank + 1, 1);
        im.put(q + 2, 1, 1);
        im.put(q + 8, s.raw, 8);
        for (int d = 0; d < rank; ++d) im.put(q + 16 + 4 * d, s.dims[d], 4);
        im.put(q + 16 + 4 * rank, 8, 4);
        q += lsize, ++nmsg;
        // modification time, version 1
        q = msg(im, q, 0x12, 8, 0);
        im.put(q, 1, 1);
        im.put(q + 4, s.mtime, 4);
        q += 8, ++nmsg;
        if (!s.extent.empty()) {
            // attribute "extent": int32 little-endian, rank-1 dataspace
            uint64_t n = s.extent.size(), dsize = pad8(4 * n);
            q = msg(im, q, 0x0C, 8 + 8 + 16 + 16 + dsize, 0);
            im.put(q, 1, 1);
            im.put(q + 2, 7, 2);
            im.put(q + 4, 12, 2);
            im.put(q + 6, 16, 2);
            im.bytes(q + 8, "extent", 6);
            const unsigned char i32[12] = {0x10, 0x08, 0, 0, 4, 0, 0, 0, 0, 0, 32, 0};
            im.bytes(q + 16, i32, 12);
            im.put(q + 32, 1, 1);
            im.put(q + 33, 1, 1);
            im.put(q + 40, n, 8);
            for (uint64_t e = 0; e < n; ++e) im.put(q + 48 + 4 * e, (uint32_t)s.extent[e], 4);
            q += 48 + dsize, ++nmsg;
        }
        if (end - q < 8) return fail("dataset header overflow");
        msg(im, q, 0x00, end - q - 8, 0);  // null message over the rest
        ++nmsg;
        im.put(h, 1, 1);
        im.put(h + 2, nmsg, 2);
        im.put(h + 4, 1, 4);
        im.put(h + 8, DSET_HDR - 16, 4);
    }
    if (std::fseek(f_, 0, SEEK_SET) != 0) return fail("seek failed");
    // the image spans [0, rawStart_), so the file always reaches its end-of-file address
    if (std::fwrite(im.b.data(), 1, im.b.size(), f_) != im.b.size()) return fail("write of the metadata failed");
    std::fflush(f_);
    return true;
}

void HDFostream::close()
{
    if (f_) std::fclose(f_);
    f_ = nullptr;
}

}  // namespace mpulse_hip
