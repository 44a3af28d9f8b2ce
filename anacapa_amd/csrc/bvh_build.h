// bvh_build.h — host BVH build (see bvh_build.cpp).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/anacapa_hip.h"

namespace ahip {

struct DNode;
struct DNode4;
struct DTri;
struct DTriShade;

// Reference-layout binary node (BVHNode, BVHBackend.h:25-39), kept for the bvh dump.
struct RefNode {
    float bmin[3];
    uint32_t dataA;
    float bmax[3];
    uint32_t dataB;
};

struct BVHOutput {
    std::vector<RefNode> refNodes;     // reference binary tree, depth-first
    std::vector<uint32_t> refOrder;    // reference m_primIndices (iota prefix + leaf order)
    std::vector<DNode> nodes;          // flattened interior nodes
    std::vector<DNode4> nodes4;        // two-level records the kernels traverse (root first)
    std::vector<DTri> tris;            // leaf-ordered triangles
    std::vector<DTriShade> shade;
    uint32_t rootRef = 0;
    float rootMin[3] = {0, 0, 0}, rootMax[3] = {0, 0, 0};
    uint32_t numTris = 0;
    uint32_t maxDepth = 0;
    uint32_t depth4 = 0;               // most DNode4 records on one root-to-leaf path
};

int buildBVH(const ahip_scene_desc* d, BVHOutput& out, std::string& err);

struct CNode;
struct DSeg;

struct CurveOutput {
    std::vector<CNode> nodes;          // median-split BVH (CurveBrute.cpp:247-320), depth-first
    std::vector<DSeg> segs;            // leaf order
    std::vector<uint32_t> segIds;      // strand, segment per record
    std::vector<float> strandInfo;     // colour rgb + meshID bits per strand
    uint32_t maxDepth = 0;
};

int buildCurves(const ahip_scene_desc* d, CurveOutput& out, std::string& err);

}  // namespace ahip
