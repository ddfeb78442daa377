// Flow-field output and restart files of the octree:
//   - Chombo-layout HDF5 (the reference's Chombo writer and reader,
//     include/iblgf/IO/chombo/chombo_impl.hpp, h5_io.hpp), encoded by h5.hpp;
//   - the tree restart file (Tree::write / Tree::read,
//     include/iblgf/domain/octree/tree.hpp:832-880).
// Fields are host arrays in the library's slot layout: component-major
// double[ncomp][nslots][n^3], n = block extent + 2, x fastest, slots in level
// then key order over the octants with data.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "octree.hpp"

namespace iblgf_amd
{

struct ChomboField
{
    std::string   name;   // component names: name (one component) or name_i
    int           ncomp;  // Field::nFields()
    const double* data;   // [ncomp][nslots][n^3]
};

// Mesh types of the reader's averaging rule (MeshObject, chombo_impl.hpp:316-339).
enum ChomboMesh
{
    kChomboCell = 0,
    kChomboFace = 1,
    kChomboEdge = 2,
    kChomboVertex = 3
};

// H5io::write_h5 (h5_io.hpp:161-173): the blocks of blocks_list_build
// (h5_io.hpp:232-263), Chombo::write_global_metaData (chombo_impl.hpp:1756-2026)
// and write_level_info (2337-2598).
std::vector<uint8_t> chombo_encode(const Octree& tr, const std::vector<ChomboField>& fields,
    bool include_correction, bool base_correction_only, double time, double dt);
void chombo_write(const Octree& tr, const std::string& path, const std::vector<ChomboField>& fields,
    bool include_correction, bool base_correction_only, double time, double dt);

// H5io::read_h5 (h5_io.hpp:96-108) + Chombo::read_u (chombo_impl.hpp:220-626):
// clears the field on the blocks of blocks_list_build(true, true) and adds
// the file's values into their leaves, averaging finer file cells.
void chombo_read(const Octree& tr, const std::string& path, const std::string& name, int ncomp, int mesh_type,
    double* data);
void chombo_read_bytes(const Octree& tr, std::vector<uint8_t> bytes, const std::string& name, int ncomp,
    int mesh_type, double* data);

// Tree::write: keys of the octants with data in depth-first order, then one
// bool leaf flag per key.
void tree_write_restart(const Octree& tr, const std::string& path);
// Tree::read: the keys and leaf flags of such a file.
void tree_read_restart(const std::string& path, std::vector<uint64_t>& keys, std::vector<uint8_t>& leafs);

}  // namespace iblgf_amd
