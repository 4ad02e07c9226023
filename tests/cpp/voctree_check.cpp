// Drives V3D_CUDA::VocabularyTreeTraversal through include/CUDA/v3d_cudamatching.h
// as a caller of the reference class would: allocate, setNodes, traverseTree
// in batches.  Usage: voctree_check <tree file> <keys file> <n keys> <out file>
// (raw float32 node descriptors and keys, int32 leaf ids out), tree shape
// branch factor 50, depth 3, allocate(3999) (not a multiple of 32: the
// reference's rounding warning, 3968 keys per batch).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CUDA/v3d_cudamatching.h"

int main(int argc, char **argv)
{
   if (argc != 5) return 2;
   typedef V3D_CUDA::VocabularyTreeTraversal<50, 3> Traversal;
   std::vector<float> nodes((size_t)Traversal::getNodeCount() * 128);
   FILE *f = std::fopen(argv[1], "rb");
   if (!f || std::fread(nodes.data(), sizeof(float), nodes.size(), f) != nodes.size()) return 3;
   std::fclose(f);
   const int n = std::atoi(argv[3]);
   std::vector<float> keys((size_t)n * 128);
   f = std::fopen(argv[2], "rb");
   if (!f || std::fread(keys.data(), sizeof(float), keys.size(), f) != keys.size()) return 4;
   std::fclose(f);
   std::vector<int> leaves((size_t)n, -7);
   try
   {
      Traversal t;
      t.allocate(3999); // rounded down to 3968 with the reference's warning
      t.setNodes(nodes.data());
      for (int i = 0; i < n; i += 3968)
      {
         const int m = n - i < 3968 ? n - i : 3968;
         t.traverseTree(m, keys.data() + (size_t)i * 128, leaves.data() + i);
      }
      t.deallocate();
   }
   catch (std::exception const &e)
   {
      std::fprintf(stderr, "%s\n", e.what());
      return 1;
   }
   f = std::fopen(argv[4], "wb");
   if (!f || std::fwrite(leaves.data(), sizeof(int), leaves.size(), f) != leaves.size()) return 5;
   std::fclose(f);
   std::printf("leaves %d (node count %d, leaf count %d)\n", n, Traversal::getNodeCount(), Traversal::getLeaveCount());
   return 0;
}
