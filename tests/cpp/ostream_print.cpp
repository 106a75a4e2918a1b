// Prints each argument, parsed as a double, the way a default-formatted std::ostream does (the
// reference writes meta.json and probe.csv through `stream << double`). Host code only.
#include <cstdlib>
#include <iostream>

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) std::cout << std::strtod(argv[a], nullptr) << '\n';
  return 0;
}
