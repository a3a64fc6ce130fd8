// Test infrastructure: a driver around the reference's own LZ4 (lz4/lz4.cc
// of the sdhash tree, built from where it lies into oracle/_ref/lz4check by
// oracle/Makefile), the compressor and decompressor that the reference's
// index files go through (sdbf/bloom_filter.cc:176-200).
//
//   lz4check c <in> <out>           LZ4_compress_limitedOutput of the file
//   lz4check d <in> <out> <size>    LZ4_uncompress of the file to size bytes
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lz4.h"

static std::vector<char> slurp(const char *p)
{
    std::vector<char> v;
    FILE *f = fopen(p, "rb");
    if (!f)
        exit(3);
    char b[1 << 16];
    size_t n;
    while ((n = fread(b, 1, sizeof b, f)) > 0)
        v.insert(v.end(), b, b + n);
    fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 4)
        return 2;
    const std::vector<char> in = slurp(argv[2]);
    std::vector<char> out;
    if (argv[1][0] == 'c') {
        out.resize(in.size() + in.size() / 255 + 64);
        const int r = LZ4_compress_limitedOutput(in.data(), out.data(), (int)in.size(), (int)out.size());
        if (r <= 0)
            return 4;
        out.resize((size_t)r);
    } else {
        if (argc < 5)
            return 2;
        out.resize((size_t)atoll(argv[4]));
        const int r = LZ4_uncompress(in.data(), out.data(), (int)out.size());
        if (r != (int)in.size())
            return 5;
    }
    FILE *f = fopen(argv[3], "wb");
    if (!f || fwrite(out.data(), 1, out.size(), f) != out.size())
        return 6;
    fclose(f);
    return 0;
}
