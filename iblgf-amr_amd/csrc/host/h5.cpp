// Minimal HDF5 encoder/decoder (superblock 0, version-1 object headers,
// symbol-table groups, contiguous datasets, version-1 attributes).  The byte
// layouts follow the published HDF5 File Format Specification (version 2.0
// of the document, sections II-IV); the defaults (group leaf K 4, internal K
// 16, 8-byte offsets and lengths, local-heap free-list sentinel 1, fill value
// message version 2) are those the HDF5 library uses when a file is created
// with its earliest-format settings, which is how the reference's
// h5_file.hpp creates its files (H5Fcreate with default property lists).
#include "h5.hpp"

#include <cstring>
#include <fstream>
#include <stdexcept>

namespace h5
{
namespace
{
constexpr uint64_t kUndef = ~0ULL;
constexpr int      kLeafK = 4;       // symbol-table node: 2K entries
constexpr int      kInternalK = 16;  // group B-tree node: 2K children
constexpr uint32_t kSnodBytes = 8 + 2 * kLeafK * 40;
constexpr uint32_t kBtreeBytes = 24 + (2 * kInternalK + 1) * 8 + 2 * kInternalK * 8;

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

struct Buf
{
    std::vector<uint8_t> b;
    size_t               size() const { return b.size(); }
    void                 u8(unsigned v) { b.push_back(static_cast<uint8_t>(v)); }
    void                 u16(unsigned v)
    {
        u8(v & 0xff);
        u8((v >> 8) & 0xff);
    }
    void u32(uint32_t v)
    {
        for (int i = 0; i < 4; ++i) u8((v >> (8 * i)) & 0xff);
    }
    void u64(uint64_t v)
    {
        for (int i = 0; i < 8; ++i) u8(static_cast<unsigned>((v >> (8 * i)) & 0xff));
    }
    void bytes(const void* p, size_t n)
    {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        b.insert(b.end(), c, c + n);
    }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void align8() { zeros(pad8(b.size()) - b.size()); }
    void put_u64(size_t at, uint64_t v)
    {
        for (int i = 0; i < 8; ++i) b[at + i] = static_cast<uint8_t>((v >> (8 * i)) & 0xff);
    }
    void put_u32(size_t at, uint32_t v)
    {
        for (int i = 0; i < 4; ++i) b[at + i] = static_cast<uint8_t>((v >> (8 * i)) & 0xff);
    }
};

void encode_type(Buf& o, const Type& t)
{
    switch (t.cls)
    {
    case Type::Int32:
        o.u8(0x10);  // version 1, class 0 (fixed point)
        o.u8(0x08);  // little endian, signed
        o.u8(0);
        o.u8(0);
        o.u32(4);
        o.u16(0);   // bit offset
        o.u16(32);  // precision
        break;
    case Type::Float64:
        o.u8(0x11);  // version 1, class 1 (floating point)
        o.u8(0x20);  // little endian, implied leading mantissa bit
        o.u8(63);    // sign bit position
        o.u8(0);
        o.u32(8);
        o.u16(0);
        o.u16(64);
        o.u8(52);  // exponent location
        o.u8(11);  // exponent size
        o.u8(0);   // mantissa location
        o.u8(52);  // mantissa size
        o.u32(1023);
        break;
    case Type::String:
        o.u8(0x13);  // version 1, class 3 (string)
        o.u8(0x00);  // null terminated, ASCII
        o.u8(0);
        o.u8(0);
        o.u32(t.size);
        break;
    case Type::Compound:
        o.u8(0x16);  // version 1, class 6 (compound)
        o.u16(static_cast<unsigned>(t.members.size()));
        o.u8(0);
        o.u32(t.size);
        for (const Member& m : t.members)
        {
            o.bytes(m.name.c_str(), m.name.size() + 1);
            o.zeros(pad8(m.name.size() + 1) - (m.name.size() + 1));
            o.u32(m.offset);
            o.u8(0);  // dimensionality (version 1 members carry array fields)
            o.zeros(3);
            o.u32(0);  // dimension permutation
            o.u32(0);
            o.zeros(16);  // dimension sizes
            encode_type(o, m.type);
        }
        break;
    default: throw std::runtime_error("h5: cannot encode datatype class");
    }
}

void encode_space(Buf& o, const Space& s)
{
    const unsigned rank = s.scalar ? 0u : static_cast<unsigned>(s.dims.size());
    o.u8(1);  // version 1
    o.u8(rank);
    o.u8(rank ? 1 : 0);  // maximum dimensions present
    o.u8(0);
    o.u32(0);
    for (unsigned i = 0; i < rank; ++i) o.u64(s.dims[i]);
    for (unsigned i = 0; i < rank; ++i) o.u64(s.dims[i]);
}

struct Message
{
    int                  type;
    int                  flags;
    std::vector<uint8_t> body;
};

Message attribute_message(const Attribute& a)
{
    Buf t, s;
    encode_type(t, a.type);
    encode_space(s, a.space);
    Buf m;
    m.u8(1);  // version 1
    m.u8(0);
    m.u16(static_cast<unsigned>(a.name.size() + 1));
    m.u16(static_cast<unsigned>(t.size()));
    m.u16(static_cast<unsigned>(s.size()));
    m.bytes(a.name.c_str(), a.name.size() + 1);
    m.zeros(pad8(a.name.size() + 1) - (a.name.size() + 1));
    m.bytes(t.b.data(), t.size());
    m.zeros(pad8(t.size()) - t.size());
    m.bytes(s.b.data(), s.size());
    m.zeros(pad8(s.size()) - s.size());
    m.bytes(a.data.data(), a.data.size());
    return Message{0x0C, 0, std::move(m.b)};
}

uint64_t emit_header(Buf& o, const std::vector<Message>& msgs)
{
    o.align8();
    const uint64_t at = o.size();
    uint32_t       hsize = 0;
    for (const auto& m : msgs) hsize += 8 + static_cast<uint32_t>(pad8(m.body.size()));
    o.u8(1);  // version 1
    o.u8(0);
    o.u16(static_cast<unsigned>(msgs.size()));
    o.u32(1);  // reference count
    o.u32(hsize);
    o.u32(0);  // alignment of the message area
    for (const auto& m : msgs)
    {
        if (m.body.size() > 0xffff) throw std::runtime_error("h5: header message too large");
        o.u16(static_cast<unsigned>(m.type));
        o.u16(static_cast<unsigned>(pad8(m.body.size())));
        o.u8(static_cast<unsigned>(m.flags));
        o.zeros(3);
        o.bytes(m.body.data(), m.body.size());
        o.zeros(pad8(m.body.size()) - m.body.size());
    }
    return at;
}
}  // namespace

// ---------------------------------------------------------------------------
// Writer
// ---------------------------------------------------------------------------
Writer::Writer()
{
    Obj r;
    r.is_group = true;
    objs_.push_back(std::move(r));
}

int Writer::add_child(int parent, const std::string& name, Obj o)
{
    if (parent < 0 || parent >= static_cast<int>(objs_.size()) || !objs_[parent].is_group)
        throw std::invalid_argument("h5: parent is not a group");
    if (name.empty() || name.find('/') != std::string::npos) throw std::invalid_argument("h5: bad object name");
    if (objs_[parent].children.count(name)) throw std::invalid_argument("h5: duplicate object name " + name);
    o.name = name;
    objs_.push_back(std::move(o));
    const int id = static_cast<int>(objs_.size()) - 1;
    objs_[parent].children[name] = id;
    return id;
}

int Writer::group(int parent, const std::string& name)
{
    Obj g;
    g.is_group = true;
    return add_child(parent, name, std::move(g));
}

int Writer::dataset(int parent, const std::string& name, const Type& t, const Space& s, const void* data,
    size_t bytes)
{
    if (bytes != s.count() * t.size) throw std::invalid_argument("h5: dataset size does not match its space");
    Obj d;
    d.type = t;
    d.space = s;
    d.data.assign(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + bytes);
    return add_child(parent, name, std::move(d));
}

void Writer::attribute(int obj, const std::string& name, const Type& t, const Space& s, const void* data,
    size_t bytes)
{
    const uint64_t n = s.scalar ? 1 : s.count();
    if (bytes != n * t.size) throw std::invalid_argument("h5: attribute size does not match its space");
    Attribute a;
    a.name = name;
    a.type = t;
    a.space = s;
    a.data.assign(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + bytes);
    objs_.at(obj).attrs.push_back(std::move(a));
}

void Writer::attr_int(int obj, const std::string& name, int v)
{
    attribute(obj, name, Type::i32(), Space::one_d(1), &v, 4);
}
void Writer::attr_double(int obj, const std::string& name, double v)
{
    attribute(obj, name, Type::f64(), Space::one_d(1), &v, 8);
}
void Writer::attr_ints(int obj, const std::string& name, const std::vector<int>& v)
{
    attribute(obj, name, Type::i32(), Space::one_d(v.size()), v.data(), 4 * v.size());
}
void Writer::attr_string(int obj, const std::string& name, const std::string& v)
{
    attribute(obj, name, Type::str(static_cast<uint32_t>(v.size() + 1)), Space::scalar_space(), v.c_str(),
        v.size() + 1);
}

std::vector<uint8_t> Writer::encode() const
{
    Buf o;
    o.zeros(96);  // superblock, written last
    std::vector<uint64_t> header(objs_.size(), kUndef), btree(objs_.size(), kUndef), heap(objs_.size(), kUndef);

    // Children before parents, so every symbol-table entry knows its header.
    std::vector<int> order;
    {
        std::vector<std::pair<int, bool>> st{{0, false}};
        while (!st.empty())
        {
            auto [id, done] = st.back();
            st.pop_back();
            if (done)
            {
                order.push_back(id);
                continue;
            }
            st.push_back({id, true});
            for (auto it = objs_[id].children.rbegin(); it != objs_[id].children.rend(); ++it)
                st.push_back({it->second, false});
        }
    }

    for (int id : order)
    {
        const Obj&           ob = objs_[id];
        std::vector<Message> msgs;
        if (ob.is_group)
        {
            // Local heap: the empty name at offset 0, then the children's names.
            std::vector<std::pair<std::string, int>> kids(ob.children.begin(), ob.children.end());
            if (kids.size() > size_t(2 * kLeafK * 2 * kInternalK))
                throw std::runtime_error("h5: group with more children than one B-tree node indexes");
            Buf                   seg;
            std::vector<uint64_t> name_off;
            seg.zeros(8);
            for (auto& k : kids)
            {
                name_off.push_back(seg.size());
                seg.bytes(k.first.c_str(), k.first.size() + 1);
                seg.align8();
            }
            o.align8();
            heap[id] = o.size();
            o.bytes("HEAP", 4);
            o.u8(0);
            o.zeros(3);
            o.u64(seg.size());
            o.u64(1);  // no free block (the library's sentinel)
            o.u64(heap[id] + 32);
            o.bytes(seg.b.data(), seg.size());

            // Symbol-table nodes of at most 2K entries each.
            std::vector<uint64_t> snod;
            std::vector<uint64_t> last_key;
            for (size_t s = 0; s < kids.size(); s += 2 * kLeafK)
            {
                const size_t n = std::min(kids.size() - s, size_t(2 * kLeafK));
                o.align8();
                snod.push_back(o.size());
                const size_t start = o.size();
                o.bytes("SNOD", 4);
                o.u8(1);
                o.u8(0);
                o.u16(static_cast<unsigned>(n));
                for (size_t i = 0; i < n; ++i)
                {
                    const int c = kids[s + i].second;
                    o.u64(name_off[s + i]);
                    o.u64(header[c]);
                    if (objs_[c].is_group)
                    {
                        o.u32(1);
                        o.u32(0);
                        o.u64(btree[c]);
                        o.u64(heap[c]);
                    }
                    else
                    {
                        o.u32(0);
                        o.u32(0);
                        o.zeros(16);
                    }
                }
                o.zeros(start + kSnodBytes - o.size());
                last_key.push_back(name_off[s + n - 1]);
            }

            o.align8();
            btree[id] = o.size();
            const size_t start = o.size();
            o.bytes("TREE", 4);
            o.u8(0);  // group node
            o.u8(0);  // leaf level
            o.u16(static_cast<unsigned>(snod.size()));
            o.u64(kUndef);
            o.u64(kUndef);
            o.u64(0);  // key 0: the empty name
            for (size_t i = 0; i < snod.size(); ++i)
            {
                o.u64(snod[i]);
                o.u64(last_key[i]);
            }
            o.zeros(start + kBtreeBytes - o.size());

            Buf st;
            st.u64(btree[id]);
            st.u64(heap[id]);
            msgs.push_back(Message{0x11, 0, st.b});
        }
        else
        {
            uint64_t daddr = kUndef;
            if (!ob.data.empty())
            {
                o.align8();
                daddr = o.size();
                o.bytes(ob.data.data(), ob.data.size());
            }
            Buf sp, ty, fv, lay;
            encode_space(sp, ob.space);
            encode_type(ty, ob.type);
            const uint8_t fill[8] = {2, 1, 2, 1, 0, 0, 0, 0};  // version 2, early allocation, default fill
            fv.bytes(fill, 8);
            lay.u8(3);  // layout message version 3
            lay.u8(1);  // contiguous
            lay.u64(daddr);
            lay.u64(ob.data.size());
            msgs.push_back(Message{0x01, 0, sp.b});
            msgs.push_back(Message{0x03, 1, ty.b});
            msgs.push_back(Message{0x05, 1, fv.b});
            msgs.push_back(Message{0x08, 0, lay.b});
        }
        for (const auto& a : ob.attrs) msgs.push_back(attribute_message(a));
        header[id] = emit_header(o, msgs);
    }

    // Superblock version 0.
    Buf s;
    const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    s.bytes(sig, 8);
    s.u8(0);  // superblock version
    s.u8(0);  // free-space storage version
    s.u8(0);  // root group symbol table entry version
    s.u8(0);
    s.u8(0);  // shared header message format version
    s.u8(8);  // size of offsets
    s.u8(8);  // size of lengths
    s.u8(0);
    s.u16(kLeafK);
    s.u16(kInternalK);
    s.u32(0);  // consistency flags
    s.u64(0);  // base address
    s.u64(kUndef);
    s.u64(o.size());  // end of file
    s.u64(kUndef);    // driver information block
    s.u64(0);         // root entry: link name offset
    s.u64(header[0]);
    s.u32(1);  // cache type: scratch pad holds the B-tree and heap
    s.u32(0);
    s.u64(btree[0]);
    s.u64(heap[0]);
    std::memcpy(o.b.data(), s.b.data(), 96);
    return std::move(o.b);
}

void Writer::write(const std::string& path) const
{
    const auto    bytes = encode();
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("h5: could not open " + path + " for writing");
    f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    if (!f) throw std::runtime_error("h5: could not write " + path);
}

// ---------------------------------------------------------------------------
// Reader
// ---------------------------------------------------------------------------
File::File(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("h5: could not open " + path);
    d_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    parse_superblock();
}

File::File(std::vector<uint8_t> bytes) : d_(std::move(bytes)) { parse_superblock(); }

void File::need(uint64_t p, uint64_t n) const
{
    if (p > d_.size() || n > d_.size() - p) throw std::runtime_error("h5: truncated or malformed file");
}
uint16_t File::u16(uint64_t p) const
{
    need(p, 2);
    return static_cast<uint16_t>(d_[p] | (d_[p + 1] << 8));
}
uint32_t File::u32(uint64_t p) const
{
    need(p, 4);
    uint32_t v = 0;
    for (int i = 3; i >= 0; --i) v = (v << 8) | d_[p + i];
    return v;
}
uint64_t File::u64(uint64_t p) const
{
    need(p, 8);
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | d_[p + i];
    return v;
}

void File::parse_superblock()
{
    const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    need(0, 96);
    if (std::memcmp(d_.data(), sig, 8) != 0) throw std::runtime_error("h5: not an HDF5 file");
    if (d_[8] != 0 && d_[8] != 1) throw std::runtime_error("h5: superblock version not supported");
    if (d_[13] != 8 || d_[14] != 8) throw std::runtime_error("h5: only 8-byte offsets and lengths supported");
    const uint64_t entry = d_[8] == 0 ? 56 : 60;  // version 1 adds 4 bytes (indexed storage K)
    root_ = u64(entry + 8);
}

std::vector<File::Msg> File::messages(uint64_t addr) const
{
    need(addr, 16);
    if (d_[addr] != 1) throw std::runtime_error("h5: only version-1 object headers supported");
    const unsigned nmsg = u16(addr + 2);
    const uint32_t hsize = u32(addr + 8);
    std::vector<std::pair<uint64_t, uint64_t>> chunks{{addr + 16, hsize}};
    std::vector<Msg>                            out;
    for (size_t c = 0; c < chunks.size() && out.size() < nmsg; ++c)
    {
        uint64_t       p = chunks[c].first;
        const uint64_t end = chunks[c].first + chunks[c].second;
        need(chunks[c].first, chunks[c].second);
        while (p + 8 <= end && out.size() < nmsg)
        {
            const int      type = u16(p);
            const uint32_t size = u16(p + 2);
            const int      flags = d_[p + 4];
            need(p + 8, size);
            if (type == 0x10) chunks.push_back({u64(p + 8), u64(p + 16)});
            if ((flags & 0x02) && type != 0x10) throw std::runtime_error("h5: shared messages not supported");
            out.push_back(Msg{type, p + 8, size});
            p += 8 + size;
        }
    }
    return out;
}

void File::btree_walk(uint64_t addr, uint64_t heap_data, std::map<std::string, uint64_t>& out) const
{
    need(addr, 24);
    if (std::memcmp(&d_[addr], "TREE", 4) != 0) throw std::runtime_error("h5: B-tree node expected");
    if (d_[addr + 4] != 0) throw std::runtime_error("h5: group B-tree expected");
    const int      level = d_[addr + 5];
    const unsigned used = u16(addr + 6);
    uint64_t       p = addr + 24 + 8;  // first key
    for (unsigned i = 0; i < used; ++i, p += 16)
    {
        const uint64_t child = u64(p);
        if (level > 0)
        {
            btree_walk(child, heap_data, out);
            continue;
        }
        need(child, 8);
        if (std::memcmp(&d_[child], "SNOD", 4) != 0) throw std::runtime_error("h5: symbol-table node expected");
        const unsigned n = u16(child + 6);
        for (unsigned e = 0; e < n; ++e)
        {
            const uint64_t ent = child + 8 + 40ULL * e;
            const uint64_t noff = u64(ent);
            const uint64_t oh = u64(ent + 8);
            uint64_t       q = heap_data + noff;
            need(q, 1);
            uint64_t end = q;
            while (end < d_.size() && d_[end]) ++end;
            out[std::string(reinterpret_cast<const char*>(&d_[q]), end - q)] = oh;
        }
    }
}

std::map<std::string, uint64_t> File::children(uint64_t group) const
{
    for (const Msg& m : messages(group))
    {
        if (m.type != 0x11) continue;
        const uint64_t bt = u64(m.pos), hp = u64(m.pos + 8);
        need(hp, 32);
        if (std::memcmp(&d_[hp], "HEAP", 4) != 0) throw std::runtime_error("h5: local heap expected");
        std::map<std::string, uint64_t> out;
        btree_walk(bt, u64(hp + 24), out);
        return out;
    }
    throw std::runtime_error("h5: object is not a symbol-table group");
}

bool File::has_child(uint64_t group, const std::string& name) const { return children(group).count(name) != 0; }

uint64_t File::child(uint64_t group, const std::string& name) const
{
    const auto c = children(group);
    auto       it = c.find(name);
    if (it == c.end()) throw std::runtime_error("h5: no object named " + name);
    return it->second;
}

Type File::decode_type(uint64_t pos, uint32_t avail, uint32_t* used) const
{
    need(pos, 8);
    if (avail < 8) throw std::runtime_error("h5: truncated datatype");
    const int      cls = d_[pos] & 0x0f;
    const int      ver = d_[pos] >> 4;
    const uint32_t size = u32(pos + 4);
    Type           t;
    t.size = size;
    switch (cls)
    {
    case 0:
        if (size != 4) throw std::runtime_error("h5: only 32-bit integers supported");
        if (d_[pos + 1] & 0x01) throw std::runtime_error("h5: big-endian integers not supported");
        t.cls = Type::Int32;
        *used = 12;
        return t;
    case 1:
        if (size != 8) throw std::runtime_error("h5: only 64-bit floats supported");
        if (d_[pos + 1] & 0x01) throw std::runtime_error("h5: big-endian floats not supported");
        t.cls = Type::Float64;
        *used = 20;
        return t;
    case 3:
        t.cls = Type::String;
        *used = 8;
        return t;
    case 6:
    {
        t.cls = Type::Compound;
        const unsigned nm = d_[pos + 1] | (d_[pos + 2] << 8);
        uint64_t       p = pos + 8;
        for (unsigned i = 0; i < nm; ++i)
        {
            Member   m;
            uint64_t e = p;
            while (e < d_.size() && d_[e]) ++e;
            need(e, 1);
            m.name.assign(reinterpret_cast<const char*>(&d_[p]), e - p);
            if (ver == 1 || ver == 2)
                p += pad8(m.name.size() + 1);
            else
                throw std::runtime_error("h5: compound datatype version not supported");
            m.offset = u32(p);
            p += 4;
            if (ver == 1) p += 28;  // dimensionality, reserved, permutation, reserved, dimension sizes
            uint32_t mu = 0;
            m.type = decode_type(p, static_cast<uint32_t>(pos + avail - p), &mu);
            p += mu;
            t.members.push_back(std::move(m));
        }
        *used = static_cast<uint32_t>(p - pos);
        return t;
    }
    default: throw std::runtime_error("h5: datatype class not supported");
    }
}

Space File::decode_space(uint64_t pos) const
{
    need(pos, 8);
    const int ver = d_[pos];
    const int rank = d_[pos + 1];
    Space     s;
    uint64_t  p;
    if (ver == 1)
    {
        p = pos + 8;
        s.scalar = rank == 0;
    }
    else if (ver == 2)
    {
        p = pos + 4;
        const int type = d_[pos + 3];
        if (type == 2) throw std::runtime_error("h5: null dataspace");
        s.scalar = type == 0;
    }
    else
        throw std::runtime_error("h5: dataspace version not supported");
    for (int i = 0; i < rank; ++i) s.dims.push_back(u64(p + 8ULL * i));
    return s;
}

bool File::has_attribute(uint64_t obj, const std::string& name) const
{
    try
    {
        attribute(obj, name);
        return true;
    }
    catch (const std::runtime_error&)
    {
        return false;
    }
}

Attribute File::attribute(uint64_t obj, const std::string& name) const
{
    for (const Msg& m : messages(obj))
    {
        if (m.type != 0x0C) continue;
        const int ver = d_[m.pos];
        if (ver < 1 || ver > 3) throw std::runtime_error("h5: attribute message version not supported");
        const unsigned nsz = u16(m.pos + 2), tsz = u16(m.pos + 4), ssz = u16(m.pos + 6);
        uint64_t       p = m.pos + (ver == 3 ? 9 : 8);
        need(p, nsz);
        std::string an(reinterpret_cast<const char*>(&d_[p]), nsz ? nsz - 1 : 0);
        an = an.c_str();
        p += ver == 1 ? pad8(nsz) : nsz;
        if (an != name) continue;
        Attribute a;
        a.name = an;
        uint32_t used = 0;
        a.type = decode_type(p, tsz, &used);
        p += ver == 1 ? pad8(tsz) : tsz;
        a.space = decode_space(p);
        p += ver == 1 ? pad8(ssz) : ssz;
        const uint64_t n = a.space.scalar ? 1 : a.space.count();
        need(p, n * a.type.size);
        a.data.assign(d_.begin() + p, d_.begin() + p + n * a.type.size);
        return a;
    }
    throw std::runtime_error("h5: no attribute named " + name);
}

int File::attr_int(uint64_t obj, const std::string& name) const
{
    Attribute a = attribute(obj, name);
    if (a.data.size() < 4) throw std::runtime_error("h5: empty attribute " + name);
    if (a.type.cls == Type::Int32)
    {
        int v;
        std::memcpy(&v, a.data.data(), 4);
        return v;
    }
    if (a.type.cls == Type::Float64)
    {
        double v;
        std::memcpy(&v, a.data.data(), 8);
        return static_cast<int>(v);
    }
    throw std::runtime_error("h5: attribute " + name + " is not numeric");
}

double File::attr_double(uint64_t obj, const std::string& name) const
{
    Attribute a = attribute(obj, name);
    if (a.type.cls == Type::Float64 && a.data.size() >= 8)
    {
        double v;
        std::memcpy(&v, a.data.data(), 8);
        return v;
    }
    if (a.type.cls == Type::Int32 && a.data.size() >= 4)
    {
        int v;
        std::memcpy(&v, a.data.data(), 4);
        return v;
    }
    throw std::runtime_error("h5: attribute " + name + " is not numeric");
}

std::string File::attr_string(uint64_t obj, const std::string& name) const
{
    Attribute a = attribute(obj, name);
    if (a.type.cls != Type::String) throw std::runtime_error("h5: attribute " + name + " is not a string");
    std::string s(a.data.begin(), a.data.end());
    return std::string(s.c_str());
}

void File::dataset_info(uint64_t obj, Type& t, Space& s, uint64_t& addr) const
{
    bool ht = false, hs = false, hl = false;
    for (const Msg& m : messages(obj))
    {
        if (m.type == 0x03)
        {
            uint32_t used = 0;
            t = decode_type(m.pos, m.size, &used);
            ht = true;
        }
        else if (m.type == 0x01)
        {
            s = decode_space(m.pos);
            hs = true;
        }
        else if (m.type == 0x08)
        {
            if (d_[m.pos] != 3 || d_[m.pos + 1] != 1)
                throw std::runtime_error("h5: only contiguous datasets (layout version 3) supported");
            addr = u64(m.pos + 2);
            hl = true;
        }
    }
    if (!ht || !hs || !hl) throw std::runtime_error("h5: object is not a dataset");
}

std::vector<double> File::read_doubles(uint64_t obj) const
{
    Type     t;
    Space    s;
    uint64_t addr = kUndef;
    dataset_info(obj, t, s, addr);
    const uint64_t      n = s.scalar ? 1 : s.count();
    std::vector<double> out(n, 0.0);
    if (addr == kUndef) return out;
    need(addr, n * t.size);
    if (t.cls == Type::Float64)
        std::memcpy(out.data(), &d_[addr], n * 8);
    else if (t.cls == Type::Int32)
        for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<int32_t>(u32(addr + 4 * i));
    else
        throw std::runtime_error("h5: dataset is not numeric");
    return out;
}

std::vector<int> File::read_member_ints(uint64_t obj, const std::string& member) const
{
    Type     t;
    Space    s;
    uint64_t addr = kUndef;
    dataset_info(obj, t, s, addr);
    const uint64_t n = s.scalar ? 1 : s.count();
    uint32_t       off = 0;
    if (!member.empty())
    {
        if (t.cls != Type::Compound) throw std::runtime_error("h5: dataset is not a compound");
        bool found = false;
        for (const Member& m : t.members)
            if (m.name == member)
            {
                if (m.type.cls != Type::Int32) throw std::runtime_error("h5: member " + member + " is not int");
                off = m.offset;
                found = true;
            }
        if (!found) throw std::runtime_error("h5: no member named " + member);
    }
    else if (t.cls != Type::Int32)
        throw std::runtime_error("h5: dataset is not int");
    std::vector<int> out(n, 0);
    if (addr == kUndef) return out;
    need(addr, n * t.size);
    for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<int32_t>(u32(addr + i * t.size + off));
    return out;
}

}  // namespace h5
