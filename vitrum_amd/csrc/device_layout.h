// device_layout.h -- HBM layout of a packed scene for the traversal kernels.
//
// Node64 (64 bytes, one L1 line): both child boxes live in the parent, so one
// 64-byte load tests two children.  It is derived losslessly from the
// reference's 32-byte preorder nodes (packages/shared-bvh/src/strides.ts:24-31):
// every reference interior node becomes one Node64 holding its two children,
// and every BLAS/TLAS root gets a "wrapper" Node64 whose left child is the
// root and whose right child is empty, so the root box test is an ordinary
// child test.
//
//   words 0-2  left child box min      words 6-8   right child box min
//   words 3-5  left child box max      words 9-11  right child box max
//   word 12    left child ref          word 13     right child ref
//   words 14-15 reserved (0)
//
// Child ref: bit 31 clear -> interior, value = Node64 index.
//            bit 31 set   -> leaf, bits 27-30 = primitive count (1..15),
//                            bits 0-26 = first primitive (triangle index for
//                            a BLAS, instance slot for the TLAS).
//            kEmptyRef    -> no child (its box is empty: +inf..-inf).
//
// Triangle record (48 bytes): a.xyz b.x | b.yz c.xy | c.z flags 0 0, indexed
// by the global packed (BVH-reordered) triangle index.
//
// Instance record (144 bytes, TLAS permutation order): world-to-local
// columns (4 x float4), local-to-world columns (4 x float4), then
// {blas wrapper node, packed instance index, blas Node128 root,
// orientation-reversing (1) or not (0)}.  Triangle record word 9 is 1 for a
// triangle whose material is single-sided (set by the engine, which knows the
// materials), else 0.
//
// Wide node (kWideKids = K children, 8K words): the same trees with levels
// collapsed.  A wide node made from a binary node holds up to K descendants
// in depth-first order (a child is replaced by its two children, largest box
// first, while fewer than K slots are used), stored axis-major so a
// traversal step loads each coordinate of all K boxes as K/4 float4s:
//
//   rows of K words: min x, min y, min z, max x, max y, max z, child refs
//   (wide node index or leaf ref), then one reserved row (0); unused slots
//   hold kEmptyRef and an empty box.  With K = 4 this is 128 bytes.
//
// A binary node is only replaced by its children when both child boxes lie
// inside its box, so a ray that reaches a child's box test reaches it in
// the binary tree too, and the boxes of the replaced nodes need no test of
// their own (each slab-test bound of a parent is implied by the child's, as
// the slab arithmetic is monotonic in the box coordinates).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace vt {

struct Pack;

constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kEmptyRef = 0x80000000u;  // leaf with zero primitives
constexpr uint32_t kMaxLeafCountInRef = 15;
constexpr uint32_t kMaxLeafOffsetInRef = (1u << 27) - 1;
constexpr uint32_t kNode64Words = 16;
constexpr uint32_t kTriWords = 12;
constexpr uint32_t kInstWords = 36;
constexpr uint32_t kWideKids = 4;                    // children per wide node
constexpr uint32_t kNode128Words = 8 * kWideKids;    // words per wide node

struct DeviceLayout {
  std::vector<uint32_t> blas_nodes;   // Node64 words
  std::vector<uint32_t> tlas_nodes;   // Node64 words
  std::vector<float> tris;            // kTriWords per triangle
  std::vector<uint32_t> instances;    // kInstWords per instance slot
  uint32_t tlas_root = 0;             // wrapper Node64 index in tlas_nodes
  uint32_t blas_node_count = 0, tlas_node_count = 0, tri_count = 0, inst_count = 0;
  std::vector<uint32_t> blas4_nodes;  // Node128 words
  std::vector<uint32_t> tlas4_nodes;  // Node128 words
  uint32_t tlas4_root = 0;
  // Worst-case traversal stack entries (TLAS part + deepest BLAS part) of the
  // Node128 trees, and of the Node64 trees.
  uint32_t stack4_need = 0, stack2_need = 0;
  uint32_t blas_need4 = 0, blas_need2 = 0;
  // Where each reference BLAS node (the pack's 32-byte preorder nodes) lives
  // in the derived trees, for refits that recompute its box on the device:
  // its parent (kNoNode for a root), the Node64 child slot holding its box
  // (Node64 index * 2 + side), and the Node128 slot holding it (node * 4 +
  // slot, or kNoNode when the collapse replaced it by its children).
  std::vector<uint32_t> blas_parent, blas_dst64, blas_dst128;
};

constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// transformLinearOrientationSign (pt-webgpu intersectionCore.wgsl.ts:96-112)
// in binary32: true when the scaled 3x3 determinant of a column-major
// local-to-world matrix is negative (the instance flips front faces).
bool orientation_reversed(const float* l2w);

// Returns an empty string on success, else the reason the pack cannot be
// represented (e.g. a leaf larger than 15 primitives).
std::string build_device_layout(const Pack& p, DeviceLayout& out);

// The TLAS part of a layout: Node64 and Node128 trees and their roots and
// stack bounds, derived from the reference TLAS nodes alone (the instance
// records do not depend on the TLAS boxes).
struct TlasLayout {
  std::vector<uint32_t> nodes, nodes4;
  uint32_t root = 0, root4 = 0, node_count = 0, need4 = 0, need2 = 0;
};
std::string build_tlas_layout(const std::vector<uint32_t>& tlas_ref_nodes, uint32_t inst_count, TlasLayout& out);
// Installs a TLAS part into a layout (stack bounds combined with its BLAS part).
void set_tlas_layout(TlasLayout&& t, DeviceLayout& out);

}  // namespace vt
