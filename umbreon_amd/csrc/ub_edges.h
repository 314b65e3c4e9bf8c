// ub_edges.h -- host side of the object-space edge generator
// (edges/object_space_edges.cpp, edges/analytic_silhouette.cpp,
// edges/mesh_feature_edges.cpp): the analytic silhouette chords of spheres
// and cylinders, the sample points of the union-boundary clip, and the
// mesh's feature edges. The point-in-solid tests of the clip run on the
// device (k_solids_contain).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/umbreon_gpu.h"

namespace ubg {

// One silhouette chord before the clip, with its source primitive (the clip
// never tests a chord against its own primitive).
struct EdgeChord {
  float a[3], b[3];
  uint16_t group;
  int32_t srcSphere, srcCyl;
};

// Edges in emission order (ubg_edges without the ownership).
struct EdgeList {
  std::vector<float> p0, p1, color, opacity1;
  std::vector<uint16_t> group;
  // makeEdge (object_space_edges.cpp:48-64)
  void push(const float a[3], const float b[3], const ubg_edge_options& o, uint16_t grp, float op0 = 1.0f,
            float op1 = 1.0f);
  int64_t size() const { return static_cast<int64_t>(group.size()); }
};

// Every sphere's horizon ring, then every cylinder's two side lines (or the
// near cap rim end-on), flattened into chords (fromEdgeMacro primitives are
// skipped).
void analyticChords(const ubg_scene_desc* d, const ubg_camera& cam, int segments, float raise,
                    std::vector<EdgeChord>& out);

// The clip's sample points of one chord: K + 1 points a + (b - a) * i / K
// with K = max(2, ceil(len / step)); none for a zero-length chord.
int chordSamples(const EdgeChord& c, float width);
void chordPoint(const EdgeChord& c, int K, int i, float p[3]);

// The solids a clip point is tested against: the scene's spheres (x, y, z, r)
// and cylinders (p0.xyz, r / p1.xyz, 0), plus a uniform grid over their
// bounds: cellStart (nCells + 1 entries) indexes cellItems, whose entries are
// sphere i (i) or cylinder j (j | 0x80000000), each listed in every cell its
// box overlaps.
struct SolidGrid {
  std::vector<float> spheres;  // 4 per sphere
  std::vector<float> cyls;     // 8 per cylinder
  float lo[3] = {0, 0, 0};
  float inv = 1.0f;            // cells per world unit
  int dim[3] = {0, 0, 0};
  std::vector<uint32_t> cellStart, cellItems;
};
void buildSolidGrid(const ubg_scene_desc* d, SolidGrid& g);

// The mesh's feature edges (extractMeshFeatureEdges with the object-space
// parameters), appended verbatim as edges.
void meshFeatureEdges(const ubg_scene_desc* d, const ubg_camera& cam, const ubg_edge_options& o, EdgeList& out);

}  // namespace ubg
