// Drives the BGEO reader (bachelor-thesis_amd/csrc/pm_bgeo.cpp) over every
// file named on the command line, as pm_load_bgeo and a caller of
// pm_bgeo_read would: built with AddressSanitizer and UBSan by
// tests/test_bgeo_sanitized.py.  Prints "<status> <points>" per file.
#include <cstdio>
#include <vector>

#include "pmarch.h"

int pm_bgeo_read_all(const char *path, std::vector<float> &out);

int main(int argc, char **argv)
{
    for (int i = 1; i < argc; i++) {
        uint64_t n = 0;
        int rc = pm_bgeo_read(argv[i], nullptr, 0, &n);
        if (rc == PM_OK && n <= (1u << 20)) {
            std::vector<float> xyz(3 * n + 3);
            rc = pm_bgeo_read(argv[i], xyz.data(), n, &n);
        }
        std::vector<float> all;
        const int rc2 = pm_bgeo_read_all(argv[i], all);
        std::printf("%d %d %llu %zu\n", rc, rc2, (unsigned long long)n, all.size() / 3);
    }
    return 0;
}
