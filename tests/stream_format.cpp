// Formats floats as the reference's cache writer does (std::ofstream << float
// with the default stream flags, Buffer.cuh:70-80) and parses decimal strings
// as its readers do (std::stof, PBD_Basic.cuh:326-530).
//
// Usage: stream_format OUT < input, one request per line:
//   w <hex bits>   write the float with these bits
//   r <text>       parse text with std::stof, write the result's hex bits
//                  ("range" when std::stof rejects it as out of range)
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::ofstream ofs(argv[1]);
    std::string op, arg;
    while (std::cin >> op >> arg) {
        if (op == "w") {
            uint32_t b = (uint32_t)std::stoul(arg, nullptr, 16);
            float f;
            std::memcpy(&f, &b, 4);
            ofs << f << "\n";
        } else {
            try {
                float f = std::stof(arg);
                uint32_t b;
                std::memcpy(&b, &f, 4);
                ofs << std::hex << b << std::dec << "\n";
            } catch (const std::out_of_range&) {
                ofs << "range\n";
            }
        }
    }
    return ofs.good() ? 0 : 1;
}
