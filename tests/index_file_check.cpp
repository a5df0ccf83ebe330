// Test harness: reads each index file named on the command line with the
// library's reader (culsh_amd/csrc/index_file.hpp, the code HashTables::load
// runs before touching the device) and prints "<status> <message>" per file.
// Built with g++ -fsanitize=address,undefined by tests/test_index_file_sanitized.py.
#include <cstdio>
#include <string>

#include "index_file.hpp"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    FILE* fp = fopen(argv[i], "rb");
    if (!fp) {
      printf("-1 cannot open\n");
      continue;
    }
    culsh_file::IndexFile f;
    std::string err;
    const int rc = culsh_file::read_index_file(fp, argv[i], f, err);
    fclose(fp);
    printf("%d %s\n", rc, err.c_str());
  }
  return 0;
}
