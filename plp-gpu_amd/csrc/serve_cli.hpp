// serve_cli.hpp -- `plp_gpu --serve` (JSON-lines request/response server).
#pragma once

// Serves requests from stdin until it ends; returns the process exit code.
int serve_main(int max_inflight, double mu_default);
