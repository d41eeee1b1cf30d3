// Decodes every file named on the command line with the host image decoder
// and reports, per file, whether it decoded.  Built by the tests with
// AddressSanitizer and UndefinedBehaviorSanitizer on the CPU, so that a
// corrupt file that reads or writes out of bounds aborts the run.
#include <cstdio>
#include <string>
#include <vector>

#include "image_decode.hpp"

int main(int argc, char** argv) {
    int decoded = 0;
    for (int i = 1; i < argc; ++i) {
        std::FILE* f = std::fopen(argv[i], "rb");
        if (!f) return 2;
        std::vector<uint8_t> data;
        unsigned char buf[65536];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + n);
        std::fclose(f);
        std::vector<uint8_t> rgba;
        uint32_t w = 0, h = 0;
        std::string error;
        bool ok = wpt::decodeImage(data.data(), data.size(), rgba, w, h, &error);
        if (ok && rgba.size() != static_cast<size_t>(w) * h * 4) return 3;
        decoded += ok;
    }
    std::printf("%d of %d decoded\n", decoded, argc - 1);
    return 0;
}
