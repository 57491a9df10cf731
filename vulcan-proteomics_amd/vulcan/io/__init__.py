"""Run stores (`<run>.vspec.arrow`): the spectra the search path reads."""
