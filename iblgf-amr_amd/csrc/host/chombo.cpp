// Chombo-layout HDF5 output and restart of the octree's fields, and the tree
// restart file.  Reference: include/iblgf/IO/chombo/chombo_impl.hpp,
// include/iblgf/IO/chombo/h5_io.hpp, include/iblgf/domain/octree/tree.hpp:832-880.
#include "chombo.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

#include "h5.hpp"

namespace iblgf_amd
{
namespace
{
// The reference's box_compound / vec_dx structs carry six (three) std::string
// member names after their values, so the compound types it registers with
// sizeof() are 216 and 120 bytes, values at the front.  The files keep those
// element sizes; the name bytes are written as zeros.
constexpr uint32_t kBoxBytes = 6 * 4 + 6 * 32;
constexpr uint32_t kVecDxBytes = 3 * 8 + 3 * 32;

h5::Type box_type()
{
    h5::Type t{h5::Type::Compound, kBoxBytes, {}};
    const char* names[6] = {"lo_i", "lo_j", "lo_k", "hi_i", "hi_j", "hi_k"};
    for (int i = 0; i < 6; ++i) t.members.push_back(h5::Member{names[i], uint32_t(4 * i), h5::Type::i32()});
    return t;
}

h5::Type vec_dx_type()
{
    h5::Type    t{h5::Type::Compound, kVecDxBytes, {}};
    const char* names[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; ++i) t.members.push_back(h5::Member{names[i], uint32_t(8 * i), h5::Type::f64()});
    return t;
}

// A block descriptor: base and extent in cells of its refinement level.
struct Box
{
    int base[3];
    int ext[3];
    int level;
    int max(int d) const { return base[d] + ext[d] - 1; }
};

// BlockDescriptor::level_scale (include/iblgf/domain/dataFields/blockDescriptor.hpp:212-236):
// to a finer level base and extent scale up; to a coarser level base and max
// divide, rounding towards minus infinity in the reference's integer form
// (which shrinks a box that does not divide exactly).
Box level_scale(Box b, int level)
{
    const int diff = std::abs(b.level - level);
    const int f = 1 << diff;
    if (level > b.level)
        for (int d = 0; d < 3; ++d)
        {
            b.base[d] *= f;
            b.ext[d] *= f;
        }
    else
        for (int d = 0; d < 3; ++d)
        {
            int mx = b.max(d);
            b.base[d] = b.base[d] < 0 ? (b.base[d] + 1) / f - 1 : b.base[d] / f;
            mx = mx < 0 ? (mx + 1) / f - 1 : mx / f;
            b.ext[d] = mx - b.base[d] + 1;
        }
    b.level = level;
    return b;
}

// BlockDescriptor::overlap (blockDescriptor.hpp:403-420): other is first
// scaled to this box's level.
bool overlap(const Box& self, Box other, Box& out)
{
    if (other.level != self.level) other = level_scale(other, self.level);
    out = other;
    for (int d = 0; d < 3; ++d)
    {
        out.base[d] = std::max(self.base[d], other.base[d]);
        out.ext[d] = std::min(self.base[d] + self.ext[d], other.base[d] + other.ext[d]) - out.base[d];
        if (out.ext[d] < 1) return false;
    }
    return true;
}

Box block_box(const Octree& tr, int o)
{
    Box b;
    tr.block_base(o, b.base);
    for (int d = 0; d < 3; ++d) b.ext[d] = tr.block_extent;
    b.level = tr.refinement_level(o);
    return b;
}

std::vector<int> slot_of(const Octree& tr)
{
    std::vector<int> slot(tr.oct.size(), -1);
    int              k = 0;
    for (const auto& lm : tr.level_map)
        for (int o : lm) slot[o] = k++;
    return slot;
}

int nslots(const Octree& tr)
{
    int k = 0;
    for (const auto& lm : tr.level_map) k += static_cast<int>(lm.size());
    return k;
}

// H5io::blocks_list_build (h5_io.hpp:232-263): octants with data at
// refinement level >= 0 in depth-first order, corrections only with
// include_correction, and then only at refinement level 0 with
// base_correction_only.
std::vector<int> blocks_list(const Octree& tr, bool include_correction, bool base_correction_only)
{
    std::vector<int> out;
    for (int o : tr.dfs_snapshot())
    {
        const Octant& oc = tr.oct[o];
        if (!oc.has_data || tr.refinement_level(o) < 0) continue;
        if (!include_correction && oc.correction) continue;
        if (base_correction_only && oc.correction && tr.refinement_level(o) > 0) continue;
        out.push_back(o);
    }
    return out;
}

// Key::operator++ at the key's level: the next key in Morton order.
uint64_t key_next(uint64_t k)
{
    const int L = Morton::level(k);
    return k + (1ULL << (Morton::shift_to_coord + 3 * (Morton::max_level - L)));
}

struct LevelInfo
{
    Box                           probe;
    std::vector<int>              octants;
    std::vector<std::vector<int>> groups;
};

// Chombo::init (chombo_impl.hpp:95-199): blocks per refinement level, the
// level's bounding box, and groups of consecutive keys that fill a whole
// subtree (8^r blocks whose first is child 0 of child 0 ... of an ancestor).
std::map<int, LevelInfo> level_info(const Octree& tr, const std::vector<int>& blocks)
{
    std::map<int, LevelInfo> lm;
    for (int o : blocks)
    {
        const Box b = block_box(tr, o);
        auto      it = lm.find(b.level);
        if (it != lm.end())
        {
            it->second.octants.push_back(o);
            for (int d = 0; d < 3; ++d)
            {
                if (it->second.probe.base[d] > b.base[d])
                {
                    // The reference moves min (base) without keeping max; the
                    // extent is then only raised by a block beyond the max.
                    it->second.probe.base[d] = b.base[d];
                }
                if (it->second.probe.max(d) < b.max(d)) it->second.probe.ext[d] = b.max(d) - it->second.probe.base[d] + 1;
            }
        }
        else
        {
            LevelInfo l;
            l.probe = b;
            l.octants.push_back(o);
            lm.emplace(b.level, l);
        }
    }
    for (auto& kv : lm)
    {
        LevelInfo&              l = kv.second;
        const std::vector<int>& oc = l.octants;
        size_t                  track = 0;
        int                     g = -1;
        for (size_t i = 0; i < oc.size(); ++i)
        {
            if (i == track)
            {
                l.groups.emplace_back();
                ++g;
                l.groups[g].push_back(oc[i]);
                unsigned recurse = 0;
                const unsigned lim = static_cast<unsigned>(tr.oct[oc[i]].level);
                int            p = oc[i];
                while (i == track && recurse < lim)
                {
                    const size_t shift = (size_t(1) << (3 * (recurse + 1))) - 1;
                    const size_t shift_final = size_t(1) << (3 * recurse);
                    const uint64_t pk = tr.oct[p].key;
                    const int      pl = Morton::level(pk);
                    if (Morton::child_number(pk, pl) == 0 && i + shift <= oc.size() - 1)
                    {
                        bool     consecutive = true;
                        uint64_t k = tr.oct[oc[i]].key;
                        for (size_t j = 0; j <= shift; ++j, k = key_next(k))
                            if (tr.oct[oc[i + j]].key != k)
                            {
                                consecutive = false;
                                break;
                            }
                        if (consecutive)
                        {
                            ++recurse;
                            p = tr.oct[p].parent;
                        }
                        else
                            track += shift_final;
                    }
                    else
                        track += shift_final;
                }
            }
            else if (i < track)
                l.groups[g].push_back(oc[i]);
        }
    }
    return lm;
}

std::vector<std::string> component_names(const std::vector<ChomboField>& fields)
{
    std::vector<std::string> c;
    for (const auto& f : fields)
    {
        if (f.ncomp == 1)
            c.push_back(f.name);
        else
            for (int i = 0; i < f.ncomp; ++i) c.push_back(f.name + "_" + std::to_string(i));
    }
    return c;
}
}  // namespace

std::vector<uint8_t> chombo_encode(const Octree& tr, const std::vector<ChomboField>& fields,
    bool include_correction, bool base_correction_only, double time, double dt)
{
    for (const auto& f : fields)
        if (f.ncomp < 1 || !f.data || f.name.empty()) throw std::invalid_argument("chombo_write: bad field");
    const std::vector<int>         blocks = blocks_list(tr, include_correction, base_correction_only);
    const std::map<int, LevelInfo> lm = level_info(tr, blocks);
    const std::vector<std::string> comps = component_names(fields);
    const int                      ncomp = static_cast<int>(comps.size());
    const int                      n = tr.block_extent + 2;
    const int                      E = tr.block_extent;
    const size_t                   n3 = size_t(n) * n * n;
    const size_t                   ns = static_cast<size_t>(nslots(tr));
    const std::vector<int>         slot = slot_of(tr);

    h5::Writer w;
    const int  root = w.root();
    // write_global_metaData (chombo_impl.hpp:1755-2026).
    for (int i = 0; i < ncomp; ++i) w.attr_string(root, "component_" + std::to_string(i), comps[i]);
    const int num_levels = static_cast<int>(lm.size());
    w.attr_int(root, "num_levels", num_levels);
    w.attr_int(root, "num_components", ncomp);
    w.attr_string(root, "filetype", "VanillaAMRFileType");
    w.attr_int(root, "data_centering", 6);
    const int glob = w.group(root, "Chombo_global");
    w.attr_int(glob, "SpaceDim", 3);

    for (int lvl = 0; lvl < num_levels; ++lvl)
    {
        auto it = lm.find(lvl);
        if (it == lm.end()) throw std::runtime_error("chombo_write: refinement levels are not contiguous");
        const LevelInfo& l = it->second;
        const int        g_lvl = w.group(root, "level_" + std::to_string(lvl));
        w.attr_int(g_lvl, "dt", static_cast<int>(dt));
        w.attr_ints(g_lvl, "ghost", {1, 1, 1});
        const double dx = tr.dx_base / std::pow(2.0, lvl);
        w.attr_double(g_lvl, "dx", dx);
        w.attr_int(g_lvl, "ref_ratio", 2);
        w.attr_int(g_lvl, "time", static_cast<int>(time));
        const int g_attr = w.group(g_lvl, "data_attributes");
        w.attr_string(g_attr, "objectType", "CArrayBox");
        {
            std::vector<uint8_t> pd(kBoxBytes, 0);
            const int            v[6] = {l.probe.base[0], l.probe.base[1], l.probe.base[2], l.probe.max(0),
                l.probe.max(1), l.probe.max(2)};
            std::memcpy(pd.data(), v, sizeof v);
            w.attribute(g_lvl, "prob_domain", box_type(), h5::Space::one_d(1), pd.data(), pd.size());
            std::vector<uint8_t> vd(kVecDxBytes, 0);
            const double         x[3] = {dx, dx, dx};
            std::memcpy(vd.data(), x, sizeof x);
            w.attribute(g_lvl, "vec_dx", vec_dx_type(), h5::Space::one_d(1), vd.data(), vd.size());
        }

        // write_level_info (chombo_impl.hpp:2336-2598).
        std::vector<int> offsets{0};
        std::vector<uint8_t> boxes(l.groups.size() * kBoxBytes, 0);
        std::vector<double>  data;
        for (size_t g = 0; g < l.groups.size(); ++g)
        {
            const std::vector<int>& grp = l.groups[g];
            // The reference sums the extents of the level's first |group|
            // octants (l.octants[b], not the group's); blocks share one extent.
            offsets.push_back(static_cast<int>(grp.size()) * E * E * E * ncomp);
            const Box b0 = block_box(tr, grp[0]);
            const int ge = static_cast<int>(std::lround(std::cbrt(static_cast<double>(grp.size()))));
            std::vector<int> ordered(grp.size(), -1);
            int              lo[3] = {b0.base[0], b0.base[1], b0.base[2]};
            int              hi[3] = {b0.max(0), b0.max(1), b0.max(2)};
            for (int o : grp)
            {
                const Box b = block_box(tr, o);
                const int gc = (b.base[0] - b0.base[0]) / E + (b.base[1] - b0.base[1]) / E * ge +
                               (b.base[2] - b0.base[2]) / E * ge * ge;
                if (gc < 0 || gc >= static_cast<int>(ordered.size()))
                    throw std::runtime_error("chombo_write: group is not a cube of blocks");
                ordered[gc] = o;
                for (int d = 0; d < 3; ++d)
                {
                    lo[d] = std::min(lo[d], b.base[d]);
                    hi[d] = std::max(hi[d], b.max(d));
                }
            }
            const int bx[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
            std::memcpy(&boxes[g * kBoxBytes], bx, sizeof bx);
            // write_data_for_each (chombo_impl.hpp:2075-2160): per component,
            // the group's cells x fastest across its blocks.
            for (const auto& f : fields)
                for (int c = 0; c < f.ncomp; ++c)
                    for (int z = 0; z < ge; ++z)
                        for (int k = 0; k < E; ++k)
                            for (int y = 0; y < ge; ++y)
                                for (int j = 0; j < E; ++j)
                                    for (int x = 0; x < ge; ++x)
                                    {
                                        const int o = ordered[x + y * ge + z * ge * ge];
                                        if (o < 0) throw std::runtime_error("chombo_write: incomplete group");
                                        const double* blk = f.data + (size_t(c) * ns + slot[o]) * n3;
                                        const double* row = blk + size_t(1) + size_t(n) * ((j + 1) + size_t(n) * (k + 1));
                                        data.insert(data.end(), row, row + E);
                                    }
        }
        const int g_attr2 = g_attr;
        w.dataset(g_attr2, "offsets", h5::Type::i32(), h5::Space::one_d(offsets.size()), offsets.data(),
            offsets.size() * 4);
        w.dataset(g_lvl, "boxes", box_type(), h5::Space::one_d(l.groups.size()), boxes.data(), boxes.size());
        w.dataset(g_lvl, "data:datatype=0", h5::Type::f64(), h5::Space::one_d(data.size()), data.data(),
            data.size() * 8);
    }
    return w.encode();
}

void chombo_write(const Octree& tr, const std::string& path, const std::vector<ChomboField>& fields,
    bool include_correction, bool base_correction_only, double time, double dt)
{
    const auto    bytes = chombo_encode(tr, fields, include_correction, base_correction_only, time, dt);
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("chombo_write: could not open " + path);
    f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    if (!f) throw std::runtime_error("chombo_write: could not write " + path);
}

namespace
{
void chombo_read_file(const Octree& tr, const h5::File& file, const std::string& name, int nfields, int mesh_type,
    double* data)
{
    if (nfields < 1 || !data) throw std::invalid_argument("chombo_read: bad field");
    if (mesh_type < kChomboCell || mesh_type > kChomboVertex) throw std::invalid_argument("chombo_read: bad mesh type");
    const int              n = tr.block_extent + 2;
    const size_t           n3 = size_t(n) * n * n;
    const size_t           ns = static_cast<size_t>(nslots(tr));
    const std::vector<int> slot = slot_of(tr);
    const std::vector<int> blocks = blocks_list(tr, true, true);

    // Cleaning (chombo_impl.hpp:224-234).
    for (int c = 0; c < nfields; ++c)
        for (int o : blocks) std::fill_n(data + (size_t(c) * ns + slot[o]) * n3, n3, 0.0);

    const uint64_t root = file.root();
    const int      num_components = file.attr_int(root, "num_components");
    const int      num_levels = file.attr_int(root, "num_levels");
    int            component_idx = -1;
    for (int i = 0; i < num_components; ++i)
    {
        const std::string a = file.attr_string(root, "component_" + std::to_string(i));
        if ((nfields > 1 && a == name + "_0") || (nfields == 1 && a == name))
        {
            component_idx = i;
            break;
        }
    }
    // The reference falls back to component 0 when the name is absent; a
    // missing field is an error here.
    if (component_idx < 0) throw std::invalid_argument("chombo_read: no component named " + name);
    if (component_idx + nfields > num_components)
        throw std::invalid_argument("chombo_read: file holds fewer components of " + name);

    const double file_dx_base = file.attr_double(file.child(root, "level_0"), "dx");
    const double diff_f = std::log2(tr.dx_base / file_dx_base);
    if (std::fabs(std::round(diff_f) - diff_f) > 1e-10) throw std::runtime_error("chombo_read: base dx doesn't match");
    const int base_level_diff = static_cast<int>(std::round(diff_f));

    for (int l = 0; l < num_levels; ++l)
    {
        const uint64_t   lg = file.child(root, "level_" + std::to_string(l));
        const int        fake_level = l + base_level_diff;
        const uint64_t   bo = file.child(lg, "boxes");
        std::vector<int> m[6];
        const char*      names[6] = {"lo_i", "lo_j", "lo_k", "hi_i", "hi_j", "hi_k"};
        for (int i = 0; i < 6; ++i) m[i] = file.read_member_ints(bo, names[i]);
        const std::vector<double> fd = file.read_doubles(file.child(lg, "data:datatype=0"));
        long long                 box_offset = 0;
        for (size_t fb = 0; fb < m[0].size(); ++fb)
        {
            Box fbox;
            for (int d = 0; d < 3; ++d)
            {
                fbox.base[d] = m[d][fb];
                fbox.ext[d] = m[3 + d][fb] - m[d][fb] + 1;
            }
            fbox.level = fake_level;
            const long long fsize = static_cast<long long>(fbox.ext[0]) * fbox.ext[1] * fbox.ext[2];
            for (int o : blocks)
            {
                if (!tr.oct[o].leaf) continue;
                const Box b = block_box(tr, o);
                if (b.level > fake_level) continue;
                const int factor = 1 << std::abs(fake_level - b.level);
                Box       ov;
                if (!overlap(b, fbox, ov)) continue;
                const Box up = level_scale(ov, fake_level);
                const int avg[2] = {(factor - 1) / 2, factor / 2};
                const int single[2] = {0, 0};
                double*   blk0 = data + size_t(slot[o]) * n3;
                for (int fi = 0; fi < nfields; ++fi)
                {
                    const int sep = nfields > 3 ? nfields / 3 : 1;
                    int       fv[3][2];
                    for (int d = 0; d < 3; ++d)
                    {
                        fv[d][0] = avg[0];
                        fv[d][1] = avg[1];
                    }
                    const int ax = fi / sep;
                    auto      set = [&](int d, const int* v) {
                        fv[d][0] = v[0];
                        fv[d][1] = v[1];
                    };
                    if (mesh_type == kChomboFace)
                        set(ax, single);
                    else if (mesh_type == kChomboEdge)
                    {
                        for (int d = 0; d < 3; ++d) set(d, single);
                        set(ax, avg);
                    }
                    else if (mesh_type == kChomboVertex)
                        for (int d = 0; d < 3; ++d) set(d, single);
                    double total = 1.0;
                    for (int d = 0; d < 3; ++d) total *= (fv[d][1] - fv[d][0] + 1);
                    double* blk = blk0 + size_t(fi) * ns * n3;
                    for (int sk = fv[2][0]; sk <= fv[2][1]; ++sk)
                        for (int k = 0; k < ov.ext[2]; ++k)
                            for (int sj = fv[1][0]; sj <= fv[1][1]; ++sj)
                                for (int j = 0; j < ov.ext[1]; ++j)
                                    for (int si = fv[0][0]; si <= fv[0][1]; ++si)
                                    {
                                        const int f2 = up.base[2] - fbox.base[2] + k * factor + sk;
                                        const int f1 = up.base[1] - fbox.base[1] + j * factor + sj;
                                        if (f2 >= fbox.ext[2] || f2 < 0 || f1 >= fbox.ext[1] || f1 < 0) continue;
                                        const long long row = box_offset +
                                                              static_cast<long long>(component_idx + fi) * fsize +
                                                              static_cast<long long>(f2) * fbox.ext[0] * fbox.ext[1] +
                                                              static_cast<long long>(f1) * fbox.ext[0];
                                        for (int i = 0; i < ov.ext[0]; ++i)
                                        {
                                            const int f0 = up.base[0] - fbox.base[0] + i * factor + si;
                                            if (f0 >= fbox.ext[0] || f0 < 0) continue;
                                            const long long at = row + f0;
                                            if (at < 0 || at >= static_cast<long long>(fd.size()))
                                                throw std::runtime_error("chombo_read: box data beyond the dataset");
                                            const int x = i + ov.base[0] - b.base[0] + 1;
                                            const int y = j + ov.base[1] - b.base[1] + 1;
                                            const int z = k + ov.base[2] - b.base[2] + 1;
                                            blk[x + size_t(n) * (y + size_t(n) * z)] += fd[at] / total;
                                        }
                                    }
                }
            }
            box_offset += fsize * num_components;
        }
    }
}
}  // namespace

void chombo_read(const Octree& tr, const std::string& path, const std::string& name, int ncomp, int mesh_type,
    double* data)
{
    h5::File f(path);
    chombo_read_file(tr, f, name, ncomp, mesh_type, data);
}

void chombo_read_bytes(const Octree& tr, std::vector<uint8_t> bytes, const std::string& name, int ncomp,
    int mesh_type, double* data)
{
    h5::File f(std::move(bytes));
    chombo_read_file(tr, f, name, ncomp, mesh_type, data);
}

void tree_write_restart(const Octree& tr, const std::string& path)
{
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("Could not open file: " + path);
    const std::vector<int> dfs = tr.dfs_snapshot();
    for (int o : dfs)
        if (tr.oct[o].has_data) f.write(reinterpret_cast<const char*>(&tr.oct[o].key), sizeof(uint64_t));
    for (int o : dfs)
        if (tr.oct[o].has_data)
        {
            const char leaf = tr.oct[o].leaf ? 1 : 0;
            f.write(&leaf, 1);
        }
    if (!f) throw std::runtime_error("Could not write file: " + path);
}

void tree_read_restart(const std::string& path, std::vector<uint64_t>& keys, std::vector<uint8_t>& leafs)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Could not open file: " + path);
    f.seekg(0, std::ios::end);
    const auto bytes = static_cast<size_t>(f.tellg());
    f.seekg(0, std::ios::beg);
    if (bytes % 9 != 0) throw std::runtime_error("tree restart file size is not a multiple of 9: " + path);
    const size_t n = bytes / 9;
    keys.resize(n);
    leafs.resize(n);
    f.read(reinterpret_cast<char*>(keys.data()), static_cast<std::streamsize>(8 * n));
    f.read(reinterpret_cast<char*>(leafs.data()), static_cast<std::streamsize>(n));
    if (!f) throw std::runtime_error("Could not read file: " + path);
}

}  // namespace iblgf_amd
