// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitize_host.py with g++, host sources only): the typed-molecule
// builder on every .xyz given, workload generation of both kinds with the
// topology of each system, and the minima database.  argv[1] = directory of
// the type tables, argv[2..] = .xyz files.
#include "fc_host.h"
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>
#include <cmath>
using namespace fc;
int main(int argc, char** argv) {
    TypeTables T; std::string err;
    const std::string d = argv[1];
    if (!T.load(d + "/ElementTypes.dat", d + "/AtomTypes.dat", d + "/BondTypes.dat", err)) { std::printf("tables: %s\n", err.c_str()); return 1; }
    int ok = 0, bad = 0;
    for (int a = 2; a < argc; a++) {
        MolSys m; bool pbc = false; double lv[9] = {0};
        if (build_from_xyz(argv[a], T, pbc, lv, m, err)) ok++; else bad++;
    }
    std::printf("builder: %d built, %d refused\n", ok, bad);
    // workload systems (both kinds) + topology of each
    const double z[3] = {0, 0, 0};
    for (int kind = 1; kind <= 2; kind++)
        for (int i = 0; i < 64; i++) {
            MolSys m;
            if (!make_workload_system(kind, i, 10001, kind == 1 ? 16 : 8, kind == 1 ? 256 : 64, 0.1, true, z, 22.4, 22.4, m)) { std::printf("gen fail %d %d\n", kind, i); continue; }
            std::vector<int> bk((size_t)(m.natoms + m.nnode) * 4);
            make_back_neighs(m.natoms, m.nnode, m.neighs.data(), bk.data());
            std::vector<int> comp(m.natoms);
            connected_components(m.natoms, m.neighs.data(), 16, comp.data());
        }
    // database
    MinimaDB db;
    for (int i = 0; i < 40; i++) {
        MolSys m;
        make_workload_system(1, i % 8, 10001, 16, 64, 0.1, false, z, 0, 0, m);
        std::vector<int> types(m.elem.begin(), m.elem.end());
        db.add_if_new(m.natoms, m.apos.data(), types.data(), -1.0 * i, i);
    }
    std::printf("database: %zu members of %d\n", db.members.size(), db.total);
    return 0;
}
