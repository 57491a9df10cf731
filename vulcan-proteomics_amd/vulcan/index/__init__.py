"""L1 peptidoform table: digestion, decoys, variable-mod expansion, content sort (input of the search path)."""
