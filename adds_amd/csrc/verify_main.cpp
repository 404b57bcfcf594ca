// adds_verify: the reference's correctness harness (verify.py:1-57) as a native tool.
//
//   adds_verify dirA/ dirB/ graph_list
//
// For every line of graph_list the graph name is the text before the first TAB (verify.py:11); files
// dirA+name and dirB+name are compared line by line as text (verify.py:28-49), reporting "not match",
// "path1 length not match" or "file2 length not match" at the first difference; the run ends with
// "all results match!!!" or "result mismatch!!!" (verify.py:53-56).
// Deliberate differences from verify.py: a missing file is skipped (verify.py would reuse the previous
// graph's open file or raise NameError), the per-graph "match" line is printed only for a match, and the
// exit status is 1 on any mismatch so scripts need not parse stdout.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s dirA/ dirB/ graph_list\n", argv[0]);
    return 2;
  }
  std::ifstream list(argv[3]);
  if (!list) {
    fprintf(stderr, "cannot open %s\n", argv[3]);
    return 2;
  }
  bool all_match = true;
  std::string line;
  while (std::getline(list, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::string graph = line.substr(0, line.find('\t'));
    std::string f1 = std::string(argv[1]) + graph, f2 = std::string(argv[2]) + graph;
    std::cout << f1 << "\n" << f2 << "\n";
    std::ifstream a(f1), b(f2);
    if (!a) {
      std::cout << "File1 not found\n";
      all_match = false;
    }
    if (!b) {
      std::cout << "File2 not found\n";
      all_match = false;
    }
    if (!a || !b) continue;
    bool ok = true;
    std::string l1, l2;
    for (;;) {
      bool g1 = (bool)std::getline(a, l1), g2 = (bool)std::getline(b, l2);
      if (!g1 || !g2) {
        if (g1) {
          std::cout << "path1 length not match\n" << l1 << "\n";
          ok = false;
        } else if (g2) {
          std::cout << "file2 length not match\n" << l2 << "\n";
          ok = false;
        }
        break;
      }
      if (l1 != l2) {
        std::cout << "not match\n" << l1 << "\n" << l2 << "\n";
        ok = false;
        break;
      }
    }
    if (ok) std::cout << "match\n";
    all_match &= ok;
  }
  std::cout << (all_match ? "all results match!!!" : "result mismatch!!!") << std::endl;
  return all_match ? 0 : 1;
}
