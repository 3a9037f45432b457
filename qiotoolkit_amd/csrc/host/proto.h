// Input files of the reference's Runner (app/runner.cc:215-260): a JSON
// document, or a folder of protobuf problem messages.
#pragma once
#include <string>

#include "json.h"

namespace qiotk {

// The problem folder `path` (utils/proto_reader.h:38-145): messages
// <path>/<name>_0.pb, <name>_1.pb, ... (name = the last path component),
// read until the next file is missing, each a serialized QuantumUtil.Problem
// (utils/problem.proto:1-30).  The first message gives the type and version,
// every message adds terms.  Returned as the equivalent input document
// {"cost_function": {"type", "version", "terms": [{"c", "ids"}, ...]}}.
Json read_proto_folder(const std::string& path);

// An input file as the Runner reads it: a folder is a protobuf problem
// folder, anything else a JSON document.
Json read_input_file(const std::string& path);

}  // namespace qiotk
