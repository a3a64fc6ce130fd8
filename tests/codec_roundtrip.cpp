// Reads a file of sdbf lines with sdbf_set(const char*) and writes every
// digest back with sdbf::to_string (no device work).
#include <iostream>

#include "sdbf_amd.h"

int main(int argc, char **argv)
{
    if (argc != 2)
        return 2;
    try {
        sdbf_set s(argv[1]);
        for (uint64_t i = 0; i < s.size(); i++)
            std::cout << s.at((uint32_t)i)->to_string();
        std::cerr << s.size() << " " << s.filter_count() << " " << s.input_size() << std::endl;
    } catch (int e) {
        std::cerr << "error " << e << std::endl;
        return 3;
    }
    return 0;
}
