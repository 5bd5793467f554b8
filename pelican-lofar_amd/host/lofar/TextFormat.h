// TextFormat.h -- the three Qt text conversions the event writers use, on
// std::string:
//
//   smartReal(v, p)   QTextStream << double with setRealNumberPrecision(p) and
//                     the default SmartNotation: the shorter of fixed and
//                     scientific with p significant digits, trailing zeros
//                     dropped, a two-digit signed exponent -- C's "%.*g".
//                     A float is streamed as the double it converts to.
//   fixedReal(v, d)   QString::number(v, 'f', d) -- C's "%.*f".
//   arg(s, a, w, f)   QString::arg(a, w, f): every occurrence of the
//                     lowest-numbered "%n" marker (n = 1..99) in s becomes a,
//                     right-aligned in w characters padded on the left with f.
#pragma once

#include <cstdio>
#include <string>

namespace pelican {
namespace ampp {
namespace text {

inline std::string smartReal(double v, int precision)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*g", precision, v);
    return buf;
}

inline std::string fixedReal(double v, int decimals)
{
    char buf[400];
    std::snprintf(buf, sizeof buf, "%.*f", decimals, v);
    return buf;
}

inline std::string arg(const std::string& s, const std::string& a, int fieldWidth = 0, char fill = ' ')
{
    // find the lowest marker number present
    int lowest = 100;
    for (size_t i = 0; i + 1 < s.size(); ++i) {
        if (s[i] != '%' || s[i + 1] < '0' || s[i + 1] > '9') continue;
        int n = s[i + 1] - '0';
        if (i + 2 < s.size() && s[i + 2] >= '0' && s[i + 2] <= '9') n = n * 10 + (s[i + 2] - '0');
        if (n >= 1 && n < lowest) lowest = n;
    }
    if (lowest == 100) return s;
    std::string padded = a;
    if (fieldWidth > (int)a.size()) padded = std::string((size_t)fieldWidth - a.size(), fill) + a;
    std::string out;
    for (size_t i = 0; i < s.size();) {
        if (s[i] == '%' && i + 1 < s.size() && s[i + 1] >= '0' && s[i + 1] <= '9') {
            int n = s[i + 1] - '0';
            size_t len = 2;
            if (i + 2 < s.size() && s[i + 2] >= '0' && s[i + 2] <= '9') {
                n = n * 10 + (s[i + 2] - '0');
                len = 3;
            }
            if (n == lowest) {
                out += padded;
                i += len;
                continue;
            }
        }
        out += s[i++];
    }
    return out;
}

}  // namespace text
}  // namespace ampp
}  // namespace pelican
