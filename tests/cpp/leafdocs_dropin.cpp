// leafdocs_dropin.cpp -- the drop-in V3D::LeafDocuments (include/Base/v3d_voctree.h,
// with the reference's serialization header on the include path) on the shared
// scenario of oracle/ref/leafdocs_scenario.h.
//   leafdocs_dropin save <binary out> <text out>   -- insert, save, print queries
//   leafdocs_dropin load <binary in>                -- load a saved file, print queries
#include <cstring>

#include "Base/v3d_voctree.h"

#include "leafdocs_scenario.h"

int main(int argc, char **argv)
{
   using namespace leafdocs_scenario;
   if (argc == 4 && !std::strcmp(argv[1], "save"))
   {
      V3D::LeafDocuments docs(kLeaves);
      for (int d = 0; d < kDocuments; ++d) docs.insertDocument(document(d));
      V3D::serializeDataToFile(argv[2], docs, true);
      V3D::serializeDataToFile(argv[3], docs, false);
      print_queries(docs, stdout);
      return 0;
   }
   if (argc == 3 && !std::strcmp(argv[1], "load"))
   {
      V3D::LeafDocuments docs(1);
      V3D::serializeDataFromFile(argv[2], docs);
      print_queries(docs, stdout);
      return 0;
   }
   return 2;
}
