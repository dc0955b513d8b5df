// market_data.cpp -- offline market data for the hexo command line (see market_data.hpp).
#include "market_data.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace hexo_cli {

namespace {

const Json kNull{};

struct Parser {
    static constexpr int kMaxDepth = 256;  // nesting bound (the API's responses nest 3 deep)
    const std::string& s;
    size_t i = 0;
    int depth = 0;
    explicit Parser(const std::string& t) : s(t) {}
    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    std::string string() {
        if (s[i] != '"') fail("expected a string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                const char e = s[i++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) fail("bad unicode escape");
                        unsigned cp = 0;
                        for (int d = 0; d < 4; ++d) {
                            const char h = s[i + d];
                            if (!std::isxdigit((unsigned char)h)) fail("bad unicode escape");
                            cp = cp * 16 + (unsigned)(h <= '9' ? h - '0' : (h | 0x20) - 'a' + 10);
                        }
                        i += 4;
                        c = cp < 0x80 ? (char)cp : '?';  // field values here are ASCII
                        break;
                    }
                    default: c = e;
                }
            }
            out.push_back(c);
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    Json value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        if (++depth > kMaxDepth) fail("nesting too deep");
        Json v = value_body();
        --depth;
        return v;
    }
    Json value_body() {
        Json v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = Json::Object;
            ws();
            if (s[i] == '}') {
                ++i;
                return v;
            }
            while (true) {
                ws();
                std::string key = string();
                ws();
                if (s[i] != ':') fail("expected ':'");
                ++i;
                v.obj[key] = value();
                ws();
                if (s[i] == ',') {
                    ++i;
                    continue;
                }
                if (s[i] == '}') {
                    ++i;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i;
            v.kind = Json::Array;
            ws();
            if (s[i] == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (s[i] == ',') {
                    ++i;
                    continue;
                }
                if (s[i] == ']') {
                    ++i;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::String;
            v.str = string();
            return v;
        }
        if (lit("null")) return v;
        if (lit("true") || lit("false")) {
            v.kind = Json::Bool;
            v.b = s[i - 1] == 'e' && s.compare(i - 4, 4, "true") == 0;
            return v;
        }
        char* end = nullptr;
        v.num = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i) fail("unexpected character");
        i = (size_t)(end - s.c_str());
        v.kind = Json::Number;
        return v;
    }
};

std::string slurp(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

// Tradier returns a single element instead of a one-element array.
std::vector<const Json*> items(const Json& v) {
    std::vector<const Json*> out;
    if (v.kind == Json::Array)
        for (const Json& e : v.arr) out.push_back(&e);
    else if (!v.is_null())
        out.push_back(&v);
    return out;
}

std::time_t local_midnight(const std::string& date) {
    std::tm t{};
    std::istringstream ss(date);
    ss >> std::get_time(&t, "%Y-%m-%d");
    if (ss.fail()) throw std::runtime_error("bad date " + date);
    t.tm_isdst = -1;
    return std::mktime(&t);
}

const double kTradingDays = (double)(5 * (365 / 7) + 365 % 7);  // src/inc/BSM.h:9
const double kExpLowerBound = 0.1;                              // src/inc/WebAPI.h (EXP_LB)

}  // namespace

const Json& Json::operator[](const std::string& key) const {
    if (kind != Object) return kNull;
    auto it = obj.find(key);
    return it == obj.end() ? kNull : it->second;
}

Json parse_json(const std::string& text) {
    Parser p(text);
    Json v = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

unsigned days_to_date(const std::string& date, const std::string& today) {
    const std::time_t then = local_midnight(date);
    const std::time_t now = today.empty() ? std::time(nullptr) : local_midnight(today);
    const double d = std::difftime(then, now) / (60.0 * 60.0 * 24.0);
    const unsigned raw = d > 0.0 ? (unsigned)d : 0u;
    return 5u * (raw / 7u) + raw % 7u;  // working days as the reference counts them
}

MarketData read_tradier_dir(const std::string& dir, const std::string& vol_type, long vol_n, const std::string& today) {
    MarketData md;
    const Json quote = parse_json(slurp(dir + "/quote.json"));
    const Json& bid = quote["quotes"]["quote"]["bid"];
    if (bid.kind != Json::Number) throw std::runtime_error(dir + "/quote.json: no quotes.quote.bid");
    md.S = bid.num;
    const Json exps = parse_json(slurp(dir + "/expirations.json"));
    if (exps["expirations"].is_null()) throw std::runtime_error(dir + ": no valid underlying (expirations is null)");
    for (const Json* date : items(exps["expirations"]["date"])) {
        if (date->kind != Json::String) continue;
        const unsigned days = days_to_date(date->str, today);
        const double T = days / kTradingDays;
        if (T <= kExpLowerBound) continue;
        options_chain chain(days, T);
        const Json doc = parse_json(slurp(dir + "/chain_" + date->str + ".json"));
        for (const Json* o : items(doc["options"]["option"])) {
            const Json& type = (*o)["option_type"];
            const Json &strike = (*o)["strike"], &ask = (*o)["ask"], &vol = (*o)[vol_type];
            if (type.kind != Json::String || type.str != "call" || (*o)["volume"].is_null() ||
                strike.kind != Json::Number || ask.kind != Json::Number || vol.kind != Json::Number)
                continue;
            if ((long)vol.num < vol_n) continue;
            option opt{};
            opt.volume = (int64_t)vol.num;
            opt.strike = strike.num;
            opt.price = ask.num;
            opt.bid = (*o)["bid"].kind == Json::Number ? (*o)["bid"].num : 0.0;
            chain.options.push_back(opt);
            chain.min_strike = std::min(chain.min_strike, opt.strike);
            chain.max_strike = std::max(chain.max_strike, opt.strike);
        }
        md.chains.push_back(std::move(chain));
    }
    return md;
}

// ---- parameter store ----

namespace {

// The sqlite3 C API entry points used here, resolved from the system library at run time.
struct Sqlite {
    void* h = nullptr;
    int (*open)(const char*, void**) = nullptr;
    int (*close)(void*) = nullptr;
    int (*exec)(void*, const char*, void*, void*, char**) = nullptr;
    int (*prepare)(void*, const char*, int, void**, const char**) = nullptr;
    int (*bind_text)(void*, int, const char*, int, void (*)(void*)) = nullptr;
    int (*bind_double)(void*, int, double) = nullptr;
    int (*step)(void*) = nullptr;
    double (*column_double)(void*, int) = nullptr;
    int (*finalize)(void*) = nullptr;
    Sqlite() {
        h = dlopen("libsqlite3.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        open = (decltype(open))dlsym(h, "sqlite3_open");
        close = (decltype(close))dlsym(h, "sqlite3_close");
        exec = (decltype(exec))dlsym(h, "sqlite3_exec");
        prepare = (decltype(prepare))dlsym(h, "sqlite3_prepare_v2");
        bind_text = (decltype(bind_text))dlsym(h, "sqlite3_bind_text");
        bind_double = (decltype(bind_double))dlsym(h, "sqlite3_bind_double");
        step = (decltype(step))dlsym(h, "sqlite3_step");
        column_double = (decltype(column_double))dlsym(h, "sqlite3_column_double");
        finalize = (decltype(finalize))dlsym(h, "sqlite3_finalize");
    }
    bool ok() const {
        return open && close && exec && prepare && bind_text && bind_double && step && column_double && finalize;
    }
};

Sqlite& sqlite() {
    static Sqlite s;
    return s;
}

constexpr int kSqliteOk = 0, kSqliteRow = 100, kSqliteDone = 101;
void (*const kSqliteTransient)(void*) = reinterpret_cast<void (*)(void*)>(-1);

// The reference's table and statements (src/inc/DB.h:31-33).
const char kCreate[] =
    "CREATE TABLE IF NOT EXISTS HestonParameters (Name TEXT NOT NULL UNIQUE, V0 REAL NOT NULL, Vm REAL NOT NULL, "
    "Rho REAL NOT NULL, Kappa REAL NOT NULL, Sigma REAL NOT NULL, ComputedAt INTEGER NOT NULL);";
const char kFetch[] = "SELECT * FROM HestonParameters WHERE Name=?1";
const char kUpsert[] =
    "INSERT INTO HestonParameters VALUES(?1,?2,?3,?4,?5,?6,strftime('%s','now')) ON CONFLICT(Name) DO UPDATE SET "
    "V0=?2,Vm=?3,Rho=?4,Kappa=?5,Sigma=?6,ComputedAt=strftime('%s','now');";

void* open_db(const std::string& data_dir) {
    Sqlite& q = sqlite();
    void* db = nullptr;
    if (q.open((data_dir + "/HParams.db").c_str(), &db) != kSqliteOk || q.exec(db, kCreate, nullptr, nullptr, nullptr) != kSqliteOk) {
        if (db) q.close(db);
        throw std::runtime_error("cannot open " + data_dir + "/HParams.db");
    }
    return db;
}

}  // namespace

bool fetch_params(const std::string& data_dir, const std::string& sym, HParams& p) {
    Sqlite& q = sqlite();
    if (!q.ok()) {
        std::ifstream in(data_dir + "/" + sym + ".params");
        return bool(in && (in >> p.v_0 >> p.v_m >> p.rho >> p.kappa >> p.sigma));
    }
    void* db = open_db(data_dir);
    void* st = nullptr;
    bool found = false;
    if (q.prepare(db, kFetch, -1, &st, nullptr) == kSqliteOk && q.bind_text(st, 1, sym.c_str(), -1, kSqliteTransient) == kSqliteOk &&
        q.step(st) == kSqliteRow) {
        p.v_0 = q.column_double(st, 1);
        p.v_m = q.column_double(st, 2);
        p.rho = q.column_double(st, 3);
        p.kappa = q.column_double(st, 4);
        p.sigma = q.column_double(st, 5);
        found = true;
    }
    if (st) q.finalize(st);
    q.close(db);
    return found;
}

void store_params(const std::string& data_dir, const std::string& sym, const HParams& p) {
    Sqlite& q = sqlite();
    if (!q.ok()) {
        std::ofstream out(data_dir + "/" + sym + ".params");
        out << std::setprecision(17) << p.v_0 << ' ' << p.v_m << ' ' << p.rho << ' ' << p.kappa << ' ' << p.sigma << '\n';
        if (!out) throw std::runtime_error("cannot write " + data_dir + "/" + sym + ".params");
        return;
    }
    void* db = open_db(data_dir);
    void* st = nullptr;
    bool ok = q.prepare(db, kUpsert, -1, &st, nullptr) == kSqliteOk &&
              q.bind_text(st, 1, sym.c_str(), -1, kSqliteTransient) == kSqliteOk;
    const double v[5] = {p.v_0, p.v_m, p.rho, p.kappa, p.sigma};
    for (int k = 0; k < 5 && ok; ++k) ok = q.bind_double(st, k + 2, v[k]) == kSqliteOk;
    ok = ok && q.step(st) == kSqliteDone;
    if (st) q.finalize(st);
    q.close(db);
    if (!ok) throw std::runtime_error("cannot store parameters for " + sym + " in HParams.db");
}

}  // namespace hexo_cli
