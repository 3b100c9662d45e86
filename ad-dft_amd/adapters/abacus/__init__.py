"""ABACUS case directories: INPUT / KPT / STRU readers and the log parser
(reference `adapters/abacus/{io,runner}.py`)."""
