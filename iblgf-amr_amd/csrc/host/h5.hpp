// Minimal HDF5 reader and writer for the files the IBLGF output layer uses.
//
// The reference writes its flow and restart fields through the HDF5 library
// (include/iblgf/IO/chombo/h5_file.hpp) in the Chombo layout.  No HDF5
// library is available to this build, so this module encodes and decodes the
// subset of the HDF5 file format those files need, as the library writes it
// for files created with its default (earliest) format settings:
//   - superblock version 0, 8-byte offsets and lengths;
//   - version-1 object headers (continuation messages followed on read);
//   - old-style groups: symbol-table message, version-1 B-tree (type 0),
//     local heap and symbol-table nodes;
//   - contiguous datasets (data layout message version 3);
//   - attribute messages version 1;
//   - datatypes: 32-bit signed little-endian integers, IEEE little-endian
//     doubles, fixed-length null-terminated strings, and compounds of those
//     (datatype message version 1).
// Anything else met on read (chunked or compact storage, dense attribute
// storage, other datatype classes) raises std::runtime_error.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace h5
{

struct Member;

struct Type
{
    enum Class { Int32 = 0, Float64 = 1, String = 3, Compound = 6 };
    int                 cls = Int32;
    uint32_t            size = 4;
    std::vector<Member> members;  // compound only

    static Type i32() { return Type{Int32, 4, {}}; }
    static Type f64() { return Type{Float64, 8, {}}; }
    static Type str(uint32_t n) { return Type{String, n, {}}; }
};

struct Member
{
    std::string name;
    uint32_t    offset = 0;
    Type        type;
};

// A dataspace: scalar (no dims) or simple of rank dims.size().
struct Space
{
    bool                  scalar = false;
    std::vector<uint64_t> dims;
    uint64_t              count() const
    {
        uint64_t c = 1;
        for (auto d : dims) c *= d;
        return c;
    }
    static Space one_d(uint64_t n) { return Space{false, {n}}; }
    static Space scalar_space() { return Space{true, {}}; }
};

struct Attribute
{
    std::string          name;
    Type                 type;
    Space                space;
    std::vector<uint8_t> data;
};

// ---------------------------------------------------------------------------
// Writer: builds the object tree in memory, encodes it on write().
// ---------------------------------------------------------------------------
class Writer
{
  public:
    Writer();
    int  root() const { return 0; }
    int  group(int parent, const std::string& name);
    int  dataset(int parent, const std::string& name, const Type& t, const Space& s, const void* data,
         size_t bytes);
    void attribute(int obj, const std::string& name, const Type& t, const Space& s, const void* data,
        size_t bytes);

    // Convenience attributes in the forms the reference's h5_file.hpp writes.
    void attr_int(int obj, const std::string& name, int v);        // simple space of one
    void attr_double(int obj, const std::string& name, double v);  // simple space of one
    void attr_ints(int obj, const std::string& name, const std::vector<int>& v);
    void attr_string(int obj, const std::string& name, const std::string& v);  // scalar

    std::vector<uint8_t> encode() const;
    void                 write(const std::string& path) const;

  private:
    struct Obj
    {
        bool                       is_group = false;
        std::string                name;
        std::vector<Attribute>     attrs;
        std::map<std::string, int> children;  // std::map orders names as strcmp does
        Type                       type;
        Space                      space;
        std::vector<uint8_t>       data;
    };
    std::vector<Obj> objs_;
    int              add_child(int parent, const std::string& name, Obj o);
};

// ---------------------------------------------------------------------------
// Reader: the whole file in memory; groups are resolved by name.
// ---------------------------------------------------------------------------
class File
{
  public:
    explicit File(const std::string& path);
    explicit File(std::vector<uint8_t> bytes);

    uint64_t root() const { return root_; }
    // Child object header address by name; throws if absent.
    uint64_t child(uint64_t group, const std::string& name) const;
    bool     has_child(uint64_t group, const std::string& name) const;
    std::map<std::string, uint64_t> children(uint64_t group) const;

    bool      has_attribute(uint64_t obj, const std::string& name) const;
    Attribute attribute(uint64_t obj, const std::string& name) const;
    int         attr_int(uint64_t obj, const std::string& name) const;
    double      attr_double(uint64_t obj, const std::string& name) const;
    std::string attr_string(uint64_t obj, const std::string& name) const;

    // A contiguous dataset's type, space and the address of its raw data
    // (undefined address for an unallocated dataset, which reads as zeros).
    void dataset_info(uint64_t obj, Type& t, Space& s, uint64_t& addr) const;
    // Member `member` (or the whole element when empty) of every element,
    // converted to double / int.
    std::vector<double> read_doubles(uint64_t obj) const;
    std::vector<int>    read_member_ints(uint64_t obj, const std::string& member) const;

    const std::vector<uint8_t>& bytes() const { return d_; }

  private:
    struct Msg
    {
        int      type;
        uint64_t pos;
        uint32_t size;
    };
    std::vector<uint8_t> d_;
    uint64_t             root_ = 0;
    void                 parse_superblock();
    std::vector<Msg>     messages(uint64_t addr) const;
    void btree_walk(uint64_t addr, uint64_t heap_data, std::map<std::string, uint64_t>& out) const;
    Type     decode_type(uint64_t pos, uint32_t avail, uint32_t* used) const;
    Space    decode_space(uint64_t pos) const;
    uint64_t u64(uint64_t p) const;
    uint32_t u32(uint64_t p) const;
    uint16_t u16(uint64_t p) const;
    void     need(uint64_t p, uint64_t n) const;
};

}  // namespace h5
