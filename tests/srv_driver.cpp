// Drives sdhashsrvHandler (include/sdhash_srv.h) through a fixed sequence of
// requests, as a Thrift client would, and prints every reply between marker
// lines.
//
//   srv_driver <home> <sources> <query files for hashString...>
#include <iostream>
#include <string>
#include <vector>

#include "sdhash_srv.h"

static void show(const char *what, const std::string &s)
{
    std::cout << "== " << what << "\n" << s << (s.empty() || s.back() == '\n' ? "" : "\n") << "== end\n";
}

int main(int argc, char **argv)
{
    if (argc < 4)
        return 2;
    std::string home = argv[1], sources = argv[2];
    std::vector<std::string> q(argv + 3, argv + argc);
    sdhashsrvHandler h(1, &home, &sources);
    std::string r;
    h.setList(r, false);
    show("setList", r);
    h.setList(r, true);
    show("setList json", r);
    h.displaySet(r, 2);
    show("displaySet 2", r);
    const int32_t r1 = h.createResultID("user");
    h.compareTwo(1, 2, 1, 0, r1);
    h.displayResult(r, r1);
    show("compareTwo", r);
    h.displayResultStatus(r, r1);
    show("status", r);
    h.displayResultInfo(r, r1);
    show("info", r);
    const int32_t r2 = h.createResultID("user");
    h.compareAll(2, 1, r2);
    h.displayResult(r, r2);
    show("compareAll", r);
    h.displayResultsList(r, "user", false);
    show("resultsList", r);
    // hash the query files into a new set, saved under the home directory
    const int32_t sid = h.createHashsetID();
    h.hashString("fresh", q, 0, sid, 0);
    h.displayContents(r, sid);
    show("hashString", r);
    h.getHashsetName(r, sid);
    show("name", r);
    // search the indexes of the loaded sets, at 16 KiB blocks
    const int32_t r3 = h.createResultID("user");
    h.hashString("search", q, 16, r3, 3);
    h.displayResult(r, r3);
    show("search", r);
    std::vector<std::string> src;
    h.displaySourceList(src);
    std::string all;
    for (const auto &s : src)
        all += s + "\n";
    show("sources", all);
    h.removeResult(r1);
    h.displayResultStatus(r, r1);
    show("removed", r);
    h.shutdown();
    return h.shutting_down() ? 0 : 1;
}
