// Driver of the caller-compiled functor path (gpumultisplit_device.hpp).
// usage: functor_driver <n> [<n> ...]   exit status 0 when every case matches
#include <stdio.h>
#include <stdlib.h>

#include <stdint.h>

#include <vector>

int functor_hash_cases(const std::vector<uint32_t>& ns);
int functor_popc_cases(const std::vector<uint32_t>& ns);
int functor_wide_cases(const std::vector<uint32_t>& ns);

int main(int argc, char** argv) {
    std::vector<uint32_t> ns;
    for (int i = 1; i < argc; ++i) ns.push_back((uint32_t)strtoul(argv[i], nullptr, 0));
    if (ns.empty()) ns = {0, 1, 6143, 6145};
    int fails = functor_hash_cases(ns);
    printf("hash functors done\n");
    fails += functor_popc_cases(ns);
    printf("popcount functors done\n");
    fails += functor_wide_cases(ns);
    printf("64-bit functors done\n");
    printf(fails ? "FUNCTOR DRIVER FAILED (%d)\n" : "FUNCTOR DRIVER OK\n", fails);
    return fails ? 1 : 0;
}
