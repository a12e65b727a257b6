// Prints the scalar at a path of keys of a YAML file read by granite_hip::load_yaml_file,
// with its line breaks escaped:  yaml_harness FILE KEY [KEY ...]
#include "granite_hip/params_yaml.hpp"

#include <iostream>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    try {
        const granite_hip::YNode root = granite_hip::load_yaml_file(argv[1]);
        const granite_hip::YNode* n = &root;
        for (int a = 2; a < argc; ++a) n = &(*n)[std::string(argv[a])];
        std::string out;
        for (char c : n->as_string()) out += (c == '\n') ? std::string("\\n") : std::string(1, c);
        std::cout << "[" << out << "]\n";
    } catch (const std::exception& ex) {
        std::cerr << ex.what() << "\n";
        return 1;
    }
    return 0;
}
