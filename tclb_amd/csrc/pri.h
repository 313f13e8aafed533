// pri.h -- the reference's binary state layout (<file>_<rank>.pri).
//
// The reference keeps each snapshot as 27 "margin blocks" (src/conf.R:771-
// 1012): block (sx, sy, sz), each side 0 = below the domain, 1 = the domain,
// 2 = above it, numbered sx + 3 sy + 9 sz + 1.  Along an axis a block spans
// n_a nodes on the domain side and the field's reach on the others.  A
// field (a density, or a density's own plane) reaches min..max along each
// axis: -e for a streamed density, widened by AddField calls and by the
// autosym mirrors (src/conf.R:137-170,373-445).  Within a block the fields
// are stacked, in declaration order, each as its own SoA sub-block.  A
// node's post-collision value goes to the domain block, or, for a field that
// only ever moves one way along an axis, to the margin block on the far side
// when the node sits on the edge it leaves through (the put of
// src/LatticeAccess.inc.cpp.Rt:128-170,327-337).  saveSolution writes the
// non-empty blocks in order, as stored values (storage_t), no header
// (src/Lattice.cu.Rt:642-661,708-737).
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "model.h"

namespace tclbx {

struct PriField {
    int plane;            // < Q: distribution plane; >= Q: extra density Q + k
    int mn[3], mx[3];     // reach along x, y, z (raw, as declared and widened)
};

struct PriBlock {
    int index;            // 1..27
    int side[3];          // 0 below, 1 domain, 2 above
    size_t nodes;         // positions per field layer: the product of n_a (domain side) or 1
    std::vector<int> fields;    // fields with a non-empty layer here, in order
    std::vector<int> layers;    // that field's thickness product (its nsize)
    size_t elems;         // values in the block: nodes * sum(layers)
    size_t offset;        // values before this block in the file
};

struct PriLayout {
    std::vector<PriField> fields;
    std::vector<PriBlock> blocks;   // the non-empty blocks, in file order
    size_t elems = 0;               // values in the file
};

// the fields of a model in the reference's declaration order, with their
// reach
std::vector<PriField> pri_fields(const ModelDef* m);
PriLayout pri_layout(const ModelDef* m, int nx, int ny, int nz);

// whether the engine's stored values of this storage format are, bit for
// bit, those of a reference storage build (double; float; float-shift and
// half-shift for the models that declare the shift w_i); `why` says what
// does not match otherwise
bool pri_exact(const ModelDef* m, int storage, std::string* why);
size_t pri_bytes(const ModelDef* m, int storage, int nx, int ny, int nz);

// host conversions between the engine's planes and the reference's bytes.
// `planes`: Q planes of N stored values, plane q at planes + q * pitch
// values; `extra`: nextra planes of N doubles (written as storage_t).
// Positions of the domain block that the reference's put never writes (the
// edge a field leaves through) are written as zero and ignored on reading.
void pri_pack(const ModelDef* m, int storage, int nx, int ny, int nz, const void* planes, size_t pitch,
              const double* extra, void* out);
void pri_unpack(const ModelDef* m, int storage, int nx, int ny, int nz, const void* in, void* planes,
                size_t pitch, double* extra);

}  // namespace tclbx
