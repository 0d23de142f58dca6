// printtime.h -- appends one row of timings to ./time_record.txt and rewrites
// the Average / Max / Min summary below the rows.
//
// Same file layout as the reference's print_time
// (benchapps/common/helper_printtime.h:10-88): every value printed fixed with
// 4 decimals in a 16-wide field, one row per run, then "Average:", "Max:" and
// "Min:" lines each followed by a row of values, then an empty line.  Existing
// rows are read back up to the line that starts with 'A'.
#pragma once

#include <algorithm>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

inline void print_time(int timerCount, double* timeVal, const char* path = "./time_record.txt")
{
    std::vector<std::vector<double>> rows;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line[0] == 'A')
                break;
            std::istringstream ss(line);
            std::vector<double> r;
            double v;
            while ((int)r.size() < timerCount && ss >> v)
                r.push_back(v);
            if ((int)r.size() == timerCount)
                rows.push_back(r);
        }
    }
    rows.emplace_back(timeVal, timeVal + timerCount);
    std::vector<double> sum(timerCount, 0.0), mx(timerCount, 0.0), mn(timerCount, 1000000.0);
    for (const auto& r : rows)
        for (int j = 0; j < timerCount; ++j) {
            sum[j] += r[j];
            mx[j] = std::max(mx[j], r[j]);
            mn[j] = std::min(mn[j], r[j]);
        }
    std::ofstream out(path, std::ios::out | std::ios::trunc);
    auto row = [&](const std::vector<double>& r) {
        for (double v : r)
            out << std::fixed << std::setprecision(4) << std::setw(16) << v;
    };
    for (const auto& r : rows) {
        row(r);
        out << "\n";
    }
    for (int j = 0; j < timerCount; ++j)
        sum[j] /= (double)rows.size();
    out << "Average:\n";
    row(sum);
    out << "\nMax:\n";
    row(mx);
    out << "\nMin:\n";
    row(mn);
    out << "\n\n";
}
