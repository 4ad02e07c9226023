// leafdocs_ref.cpp -- the reference's own V3D::LeafDocuments
// (V3D/include/Base/v3d_voctree.h, header-only) on the shared scenario: writes
// its binary and text archives and prints the query results.  Built by
// oracle/Makefile (target `ref`) from the reference tree into oracle/_ref/;
// test infrastructure only.
//   leafdocs_ref <binary archive out> <text archive out>
#include "Base/v3d_voctree.h"

#include "leafdocs_scenario.h"

int main(int argc, char **argv)
{
   if (argc != 3) return 2;
   using namespace leafdocs_scenario;
   V3D::LeafDocuments docs(kLeaves);
   for (int d = 0; d < kDocuments; ++d) docs.insertDocument(document(d));
   V3D::serializeDataToFile(argv[1], docs, true);
   V3D::serializeDataToFile(argv[2], docs, false);
   print_queries(docs, stdout);
   return 0;
}
