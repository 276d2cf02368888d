// Writes records with std::ofstream the way the reference's writers configure
// their streams, as the byte-level expectation for the library's writers:
//   data record   step, then std::scientific, setprecision(10) values, each
//                 followed by a space, then std::endl
//                 (cilia_sim_main.cu:584-606, mobility_solver.cpp:3076-3153)
//   reference     default-format values, each followed by a space, then "\n"
//                 (swimmer.cpp:1114-1124)
//   .par line     value << " " << "%% NAME" << std::endl (cilia_sim_main.cu:170-185)
//
// usage: iostream_records <doubles.bin> <step> <record.dat> <refs.dat> <par.par>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 6) return 2;
    std::vector<double> v;
    {
        std::FILE* f = std::fopen(argv[1], "rb");
        if (!f) return 3;
        double x;
        while (std::fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
        std::fclose(f);
    }
    const long step = std::atol(argv[2]);
    {
        std::ofstream out(argv[3], std::ios::app);
        out << step << " ";
        out << std::scientific << std::setprecision(10);
        for (double x : v) out << x << " ";
        out << std::endl;
    }
    {
        std::ofstream out(argv[4]);
        for (double x : v) out << x << " ";
        out << "\n";
    }
    {
        std::ofstream out(argv[5]);
        out << 640 << " " << "%% NFIL" << std::endl;
        out << step << " " << "%% TOTAL_TIME_STEPS" << std::endl;
        for (size_t i = 0; i < v.size(); ++i) out << v[i] << " " << "%% P" << i << std::endl;
    }
    return 0;
}
