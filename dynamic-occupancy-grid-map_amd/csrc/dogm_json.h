// A small JSON reader (objects, arrays, strings, numbers, true / false / null) for the tuning
// table; pure C++ with no HIP dependency, so the host sanitizer build exercises it too.
#pragma once

#include <cctype>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace dogm_json
{

struct JVal
{
    enum Kind { NUL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const
    {
        for (auto& kv : obj)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
};

struct Parser
{
    const std::string& s;
    size_t i = 0;
    bool ok = true;

    void ws()
    {
        while (i < s.size() && std::isspace((unsigned char)s[i]))
            ++i;
    }
    bool eat(char c)
    {
        ws();
        if (i < s.size() && s[i] == c)
        {
            ++i;
            return true;
        }
        return false;
    }
    std::string string()
    {
        std::string out;
        if (!eat('"'))
        {
            ok = false;
            return out;
        }
        while (i < s.size() && s[i] != '"')
        {
            if (s[i] == '\\' && i + 1 < s.size())
                ++i;
            out += s[i++];
        }
        if (i >= s.size())
            ok = false;
        ++i;
        return out;
    }
    JVal value(int depth = 0)
    {
        JVal v;
        ws();
        if (!ok || i >= s.size() || depth > 16)
        {
            ok = false;
            return v;
        }
        const char c = s[i];
        if (c == '{')
        {
            ++i;
            v.kind = JVal::OBJ;
            if (eat('}'))
                return v;
            do
            {
                std::string k = string();
                if (!eat(':'))
                {
                    ok = false;
                    return v;
                }
                v.obj.emplace_back(k, value(depth + 1));
            } while (ok && eat(','));
            if (!eat('}'))
                ok = false;
        }
        else if (c == '[')
        {
            ++i;
            v.kind = JVal::ARR;
            if (eat(']'))
                return v;
            do
                v.arr.push_back(value(depth + 1));
            while (ok && eat(','));
            if (!eat(']'))
                ok = false;
        }
        else if (c == '"')
        {
            v.kind = JVal::STR;
            v.str = string();
        }
        else if (s.compare(i, 4, "null") == 0 || s.compare(i, 4, "true") == 0)
        {
            i += 4;
        }
        else if (s.compare(i, 5, "false") == 0)
        {
            i += 5;
        }
        else
        {
            char* end = nullptr;
            v.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i)
                ok = false;
            i = (size_t)(end - s.c_str());
            v.kind = JVal::NUM;
        }
        return v;
    }
};

// A JSON number as an integer clamped to [lo, hi] (NaN -> lo): no out-of-range conversion.
inline long long clamp_ll(double v, long long lo = -(1ll << 62), long long hi = 1ll << 62)
{
    if (!(v >= (double)lo))
        return lo;
    if (v >= (double)hi)
        return hi;
    return (long long)v;
}

}  // namespace dogm_json
