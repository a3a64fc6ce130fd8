"""The oracle's known-answer tests and pins must fail on plausible mistakes:
each case below plants one in a copy of oracle/sdbf_oracle.c (on the CPU,
in a temporary directory), builds it, and runs test_oracle_kat.py,
test_oracle_index_kat.py and test_oracle_pins.py against that build, which
must fail.  CPU only."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "sdbf_oracle.c")

MUTANTS = {
    # gen_chunk_hash (sdbf/sdbf_core.cc:94-136)
    "restart count kept across units": [
        ("    uint32_t bigfi_count = 0;\n    if (chunk_size <= ORC_WIN)",
         "    static uint32_t bigfi_count = 0;\n    if (chunk_size <= ORC_WIN)")],
    "big filter cleared per unit": [
        ("    uint32_t bigfi_count = 0;\n    if (chunk_size <= ORC_WIN)\n        return;",
         "    uint32_t bigfi_count = 0;\n    memset(st->big, 0, ORC_BIG);\n    if (chunk_size <= ORC_WIN)\n        return;")],
    "big filter checked before the filter": [
        ("            if (!orc_bf_sha1_insert(cur, h))\n                continue;\n"
         "            if (index && !index_query_and_set(index, h, 1))\n                continue;\n"
         "            if (!big_insert(st->big, h))\n                continue;",
         "            if (!big_insert(st->big, h))\n                continue;\n"
         "            if (!orc_bf_sha1_insert(cur, h))\n                continue;\n"
         "            if (index && !index_query_and_set(index, h, 1))\n                continue;")],
    "rollover at 161": [("if (last_count == ORC_MAX_ELEM) {", "if (last_count == ORC_MAX_ELEM + 1) {")],
    "big filter restart at 8737": [("if (bigfi_count == ORC_BIG_ELEM) {", "if (bigfi_count == ORC_BIG_ELEM - 1) {")],
    "threshold >= 16": [("if (scores[i] > ORC_THRESHOLD) {", "if (scores[i] >= ORC_THRESHOLD) {")],
    "trim at <= 20": [("st.last_count < ORC_MAX_ELEM / 8", "st.last_count <= ORC_MAX_ELEM / 8")],
    # gen_block_hash and the threshold walk (sdbf/sdbf_core.cc:142-191, 271-276, 309-319)
    "allowance spent on skipped features": [
        ("            feat_words(fs, pos + i, h);\n            if (!orc_bf_sha1_insert(bf, h))\n                continue;\n"
         "            if (nsearch == 0) {",
         "            feat_words(fs, pos + i, h);\n            if (scores[i] == threshold)\n                allowed--;\n"
         "            if (!orc_bf_sha1_insert(bf, h))\n                continue;\n            if (nsearch == 0) {")],
    "tail threshold 15": [("ORC_THRESHOLD, ORC_MAX_ELEM_DD,\n", "ORC_THRESHOLD - 1, ORC_MAX_ELEM_DD,\n")],
    "walk stops at 16": [("    for (k = 65, sum = 0; k >= ORC_THRESHOLD; k--) {",
                          "    for (k = 65, sum = 0; k > ORC_THRESHOLD; k--) {")],
    # sdbf_score (sdbf/sdbf_core.cc:331-389)
    "swap with and": [("((elem_count(s1, n1 - 1) > elem_count(s2, s2->bf_count - 1)) ||",
                       "((elem_count(s1, n1 - 1) > elem_count(s2, s2->bf_count - 1)) &&")],
    "plain sum": [("score_sum = (score_sum < 0) ? m : score_sum + m;", "score_sum = score_sum + m;")],
    "sparse always off the denominator": [("    if (n1 > 1)\n        denom -= sparse;", "    denom -= sparse;")],
    "round half to even": [("return (int)floor(v + 0.5);", "return (int)rint(v);")],
    "sampled swap reads the last count": [
        ("((elem_count(s1, n1 - 1) > elem_count(s2, s2->bf_count - 1)) ||",
         "((elem_count(s1, s1->bf_count - 1) > elem_count(s2, s2->bf_count - 1)) ||")],
    "sample ignored": [("    if (sample > 0 && s1->bf_count > sample)\n        n1 = sample;",
                        "    if (0)\n        n1 = sample;")],
    # the index paths (sdbf/sdbf_core.cc:110-115, 160-177; bloom_filter.cc:52-71, 334-359)
    "index gate after the big filter": [
        ("            if (index && !index_query_and_set(index, h, 1))\n                continue;\n"
         "            if (!big_insert(st->big, h))\n                continue;",
         "            if (!big_insert(st->big, h))\n                continue;\n"
         "            if (index && !index_query_and_set(index, h, 1))\n                continue;")],
    "index gate before the filter": [
        ("            if (!orc_bf_sha1_insert(cur, h))\n                continue;\n"
         "            if (index && !index_query_and_set(index, h, 1))\n                continue;",
         "            if (index && !index_query_and_set(index, h, 1))\n                continue;\n"
         "            if (!orc_bf_sha1_insert(cur, h))\n                continue;")],
    "every counted feature looked up": [("            } else if (hash_cnt % 4 == 0) {", "            } else {")],
    "lookup after the count": [("            } else if (hash_cnt % 4 == 0) {", "            } else if ((hash_cnt + 1) % 4 == 0) {")],
    "insert while searching": [("            if (nsearch == 0) {\n                if (info && info->index)",
                                "            if (1) {\n                if (info && info->index)")],
    "index mask one bit short": [("    uint32_t bits = log_size + 2;", "    uint32_t bits = log_size + 1;")],
    "insert counts every element": [("        if (bit_cnt < 5) {\n            ix->elem_count++;",
                                     "        ix->elem_count++;\n        if (bit_cnt < 5) {")],
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(name, tmp_path):
    src = open(SRC).read()
    for a, b in MUTANTS[name]:
        assert src.count(a) == 1, a
        src = src.replace(a, b)
    c = tmp_path / "mutant.c"
    c.write_text(src)
    so = tmp_path / "libmutant.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-I", os.path.join(ROOT, "oracle"),
                    "-o", str(so), str(c), "-lcrypto", "-lm"], check=True, cwd=os.path.join(ROOT, "oracle"))
    env = dict(os.environ, SDBF_ORACLE_SO=str(so))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_kat.py"),
                        os.path.join(ROOT, "tests", "test_oracle_index_kat.py"),
                        os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                       env=env, capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 1 and " failed" in r.stdout, r.stdout[-2000:]
