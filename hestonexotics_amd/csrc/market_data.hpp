// market_data.hpp -- offline market data for the hexo command line.
//
// The reference downloads quotes, expirations and option chains from the Tradier API
// (src/WebAPI.cpp:53-136) and keeps calibrated parameters in sqlite (src/DB.cpp,
// src/inc/DB.h:31-33). Offline, the same data comes from files:
//   * a directory $HEXO_DATA_DIR/<SYM>/ holding the API's JSON responses saved as
//       quote.json                 (QUOTE_URL: quotes.quote.bid)
//       expirations.json           (EXP_URL: expirations.date[])
//       chain_<YYYY-MM-DD>.json    (OPTIONS_CHAIN_URL: options.option[] with option_type, strike,
//                                   ask, bid, volume, open_interest)
//     read with the reference's filters (calls only; volume, strike and ask present; the selected
//     volume field >= the threshold; expiries with T <= 0.1 years dropped) and its day count
//     (src/WebAPI.cpp:22-31, relative to $HEXO_TODAY=YYYY-MM-DD or the current date);
//   * HParams.db in $HEXO_DATA_DIR with the reference's table, through the system's libsqlite3
//     loaded at run time; without it, <SYM>.params text files.
#pragma once

#include <list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hexo/Types.h"

namespace hexo_cli {

// A parsed JSON value (objects, arrays, strings, numbers, true/false/null).
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;
    const Json& operator[](const std::string& key) const;  // Null when absent or not an object
    bool is_null() const { return kind == Null; }
};
Json parse_json(const std::string& text);  // throws std::runtime_error on malformed input

struct MarketData {
    double S = 0.0;
    std::list<options_chain> chains;
};

// Working days to an expiry date as the reference counts them (src/WebAPI.cpp:22-31).
unsigned days_to_date(const std::string& date, const std::string& today);
MarketData read_tradier_dir(const std::string& dir, const std::string& vol_type, long vol_n, const std::string& today);

// Parameter store: HParams.db (reference schema) when libsqlite3 is available, else text files.
bool fetch_params(const std::string& data_dir, const std::string& sym, HParams& p);
void store_params(const std::string& data_dir, const std::string& sym, const HParams& p);

}  // namespace hexo_cli
